/*
 * accel_fold.c -- the consumer of the .cand files: prepfold -accelcand on
 * a .dat (src/prepfold_pipeline.c:740-767, 1375-1594; src/fold.c;
 * src/get_candidates.c:27-37).  A candidate's record is read, its average
 * r, z, w become the start f, fdot, fdotdot of the fold, and the series is
 * folded on the GPU (fold_gpu.hip) into npart sub-integration profiles
 * with fold()'s statistics; prepfold's period / period-derivative search
 * over those sub-integrations then runs on the GPU (accel_fold_search).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "presto_accel.h"
#include "accel_internal.h"
#include "agpu.h"

#define FOLD_REC 1024   /* SUBSBLOCKLEN (include/presto.h:58) */

int accel_read_rzw_cand(const char *candfile, int candnum, accel_fourierprops *out)
{
    if (candnum < 1) {
        accel_set_error("candidate numbers start at 1 (got %d)", candnum);
        return -1;
    }
    FILE *f = fopen(candfile, "rb");
    if (!f) {
        accel_set_error("cannot open '%s'", candfile);
        return -1;
    }
    int rc = -1;
    if (fseek(f, (long) (candnum - 1) * (long) sizeof(accel_fourierprops), SEEK_SET) == 0 &&
        fread(out, sizeof(accel_fourierprops), 1, f) == 1)
        rc = 0;
    else
        accel_set_error("'%s' has no candidate %d", candfile, candnum);
    fclose(f);
    return rc;
}

void accel_cand_fold_params(const accel_fourierprops *c, double T, double *f, double *fd,
                            double *fdd)
{
    /* the .cand holds the average r and z; the fold wants start values */
    const double z0 = c->z - 0.5 * c->w;
    const double r0 = c->r - 0.5 * z0 - c->w / 6.0;
    *f = r0 / T;
    *fd = z0 / (T * T);
    *fdd = c->w / (T * T * T);
}

int accel_fold_proflen(double f, double dt)
{
    long n = (long) (1.0 / f / dt + 0.5);
    return (int) (n > 64 ? 64 : (n < 2 ? 2 : n));
}

static double phase_at(double t, double f, double fd, double fdd)
{
    return t * (t * (t * (fdd / 6.0) + fd / 2.0) + f);
}

static void profile_stats(const double *prof, int n, double numdata, double data_avg,
                          double data_var, accel_foldstats *st)
{
    st->numdata = numdata;
    st->data_avg = data_avg;
    st->data_var = data_var;
    st->numprof = n;
    double avg = 0.0;
    for (int k = 0; k < n; k++)
        avg += prof[k];
    avg /= n;
    double chi = 0.0;
    for (int k = 0; k < n; k++)
        chi += (prof[k] - avg) * (prof[k] - avg);
    st->prof_avg = avg;
    st->prof_var = data_var * numdata / n;
    st->redchi = st->prof_var > 0.0 ? chi / (st->prof_var * (n - 1)) : 0.0;
}

long long accel_fold(const float *series, long long n, double dt, double f, double fd,
                     double fdd, int npart, int proflen, double *rawfolds,
                     accel_foldstats *parts, accel_foldstats *total, int device)
{
    if (npart < 1 || proflen < 2 || dt <= 0.0 || !(f > 0.0)) {
        accel_set_error("fold: npart %d, proflen %d, dt %g, f %g out of range", npart, proflen,
                        dt, f);
        return -1;
    }
    /* records of 1024 points, a whole number per part
     * (src/prepfold_pipeline.c:536-551) */
    const long long rpp = n / FOLD_REC / npart;
    if (rpp < 1) {
        accel_set_error("fold: %lld points make fewer than %d records of %d", n, npart,
                        FOLD_REC);
        return -1;
    }
    if (rpp > 0x7fffffffLL / npart) {
        accel_set_error("fold: too many records");
        return -1;
    }
    const long long nrec = rpp * npart, used = nrec * FOLD_REC;
    double *part_end = (double *) malloc(sizeof(double) * npart);
    double *recsums = (double *) malloc(sizeof(double) * 2 * (size_t) nrec);
    const double proftime = FOLD_REC * dt;
    const double phi0 = phase_at(0.0, f, fd, fdd);
    for (int p = 0; p < npart; p++) {
        /* end time of the part's last point: its record's start + 1024 dt */
        const double tlo = (double) ((long long) p * rpp) * proftime + (rpp - 1) * proftime;
        part_end[p] = phase_at(tlo + FOLD_REC * dt, f, fd, fdd) - phi0;
    }
    void *stream = NULL, *d_x = NULL, *d_end = NULL, *d_prof = NULL, *d_sums = NULL;
    const size_t pbytes = sizeof(double) * (size_t) npart * proflen;
    long long rc = -1;
    if (agpu_init(device) || agpu_stream_create(&stream) ||
        agpu_malloc(&d_x, sizeof(float) * (size_t) used) ||
        agpu_malloc(&d_end, sizeof(double) * npart) || agpu_malloc(&d_prof, pbytes) ||
        agpu_malloc(&d_sums, sizeof(double) * 2 * (size_t) nrec) ||
        agpu_h2d(d_x, series, sizeof(float) * (size_t) used, stream) ||
        agpu_h2d(d_end, part_end, sizeof(double) * npart, stream) ||
        agpu_memset(d_prof, 0, pbytes, stream)) {
        accel_set_error("fold: %s", agpu_last_error());
        goto done;
    }
    agpu_fold_args a;
    memset(&a, 0, sizeof(a));
    a.series = (const float *) d_x;
    a.dt = dt;
    a.f = f;
    a.fd = fd;
    a.fdd = fdd;
    a.phi_start = phi0;
    a.npart = npart;
    a.rpp = (int) rpp;
    a.proflen = proflen;
    a.part_end = (const double *) d_end;
    a.rawfolds = (double *) d_prof;
    a.recsums = (double *) d_sums;
    if (agpu_fold(&a, stream) || agpu_d2h(rawfolds, d_prof, pbytes, stream) ||
        agpu_d2h(recsums, d_sums, sizeof(double) * 2 * (size_t) nrec, stream) ||
        agpu_stream_sync(stream)) {
        accel_set_error("fold: %s", agpu_last_error());
        goto done;
    }
    /* statistics: per part, then of the summed profile */
    double all1 = 0.0, all2 = 0.0;
    double *sum = (double *) calloc(proflen, sizeof(double));
    for (int p = 0; p < npart; p++) {
        double s1 = 0.0, s2 = 0.0;
        for (long long r = (long long) p * rpp; r < (long long) (p + 1) * rpp; r++) {
            s1 += recsums[2 * r];
            s2 += recsums[2 * r + 1];
        }
        all1 += s1;
        all2 += s2;
        const double m = (double) (rpp * FOLD_REC);
        const double avg = s1 / m, var = (s2 - s1 * avg) / (m - 1.0);
        if (parts)
            profile_stats(rawfolds + (size_t) p * proflen, proflen, m, avg, var, &parts[p]);
        for (int k = 0; k < proflen; k++)
            sum[k] += rawfolds[(size_t) p * proflen + k];
    }
    if (total) {
        const double m = (double) used;
        const double avg = all1 / m, var = (all2 - all1 * avg) / (m - 1.0);
        profile_stats(sum, proflen, m, avg, var, total);
    }
    free(sum);
    rc = used;
done:
    agpu_free(d_x);
    agpu_free(d_end);
    agpu_free(d_prof);
    agpu_free(d_sums);
    if (stream)
        agpu_stream_destroy(stream);
    free(part_end);
    free(recsums);
    return rc;
}

/* src/prepfold_utils.c:7-73 */
static double pfdot_of(double pf, double pfdot)
{
    if (pf == 0.0)
        return 0.0;
    const double r = -pfdot / (pf * pf);
    return r == -0 ? 0.0 : r;
}

static double pfdotdot_of(double pf, double pfdot, double pfdotdot)
{
    if (pf == 0.0 || pfdotdot == 0.0)
        return 0.0;
    const double r = 2.0 * pfdot * pfdot / (pf * pf * pf) - pfdotdot / (pf * pf);
    return r == -0 ? 0.0 : r;
}

static double fdot_of_delay(double phasedelay, double time)
{
    return time == 0.0 ? 0.0 : 2.0 * phasedelay / (time * time);
}

/* combine_profs for one trial (src/fold.c:696-754): the best profile */
static void combine_trial(const double *profs, int npart, int proflen, const double *parttimes,
                          double fdotdot, int searchpdd, double fdot, int tot, double *out)
{
    for (int k = 0; k < proflen; k++)
        out[k] = 0.0;
    for (int ii = 0; ii < npart; ii++) {
        const double t = parttimes[ii];
        double pdd_delay = 0.0;
        if (searchpdd) {
            double pdd = fdotdot * t * t * t / 6.0;
            if (pdd == -0)
                pdd = 0.0;
            pdd_delay = pdd * proflen;
        }
        double pd = fdot * t * t / 2.0;
        if (pd == -0)
            pd = 0.0;
        const double delay = (pdd_delay + pd * proflen) + (double) (ii * tot) / npart;
        double ld = fmod(-delay, (double) proflen);
        if (ld < 0.0)
            ld += proflen;
        const int off = (int) (ld + 0.5);
        for (int jj = 0; jj < proflen; jj++) {
            int kk = jj - off;
            if (kk < 0)
                kk += proflen;
            out[kk] += profs[(size_t) ii * proflen + jj];
        }
    }
}

int accel_fold_search(const double *rawfolds, const accel_foldstats *parts, int npart,
                      int proflen, double f, double fd, double fdd, double T,
                      const double *parttimes, int npfact, int pstep, int pdstep, int nosearch,
                      int searchpdd, float *ppdot, double *bestprof, accel_fold_best *best,
                      int device)
{
    if (npart < 1 || proflen < 2 || npfact < 1 || pstep < 1 || pdstep < 1 || !(T > 0.0) ||
        !(f > 0.0) || !best) {
        accel_set_error("fold search: bad arguments");
        return -1;
    }
    /* without a search the reference keeps the plane it shows small
     * (src/prepfold_pipeline.c:1648-1654) */
    const int nt = nosearch && proflen > 100 ? 201 : 2 * npfact * proflen + 1;
    if (nt > 65535) {
        accel_set_error("fold search: %d trials per axis is too many", nt);
        return -1;
    }
    /* src/prepfold_pipeline.c:1696-1705 */
    double *periods = (double *) malloc(sizeof(double) * nt);
    double *fdots = (double *) malloc(sizeof(double) * nt);
    const int npdd = searchpdd && !nosearch ? nt : 1;
    const size_t ncube = (size_t) npdd * nt * nt;
    double *fdotdots = (double *) malloc(sizeof(double) * nt);
    double *pp = (double *) malloc(sizeof(double) * ncube);
    double *prof = (double *) malloc(sizeof(double) * proflen);
    if (!periods || !fdots || !fdotdots || !pp || !prof) {
        free(periods);
        free(fdots);
        free(fdotdots);
        free(pp);
        free(prof);
        accel_set_error("fold search: out of host memory");
        return -1;
    }
    for (int i = 0; i < nt; i++) {
        const int tot = i - (nt - 1) / 2;
        double d = (double) (tot * pstep) / proflen;
        periods[i] = 1.0 / (f + d / T);
        d = (double) (tot * pdstep) / proflen;
        fdots[i] = fdot_of_delay(d, T);
        fdotdots[i] = 6.0 * d / (T * T * T);   /* phasedelay2fdotdot, T > 0 */
    }
    accel_foldstats st;
    memset(&st, 0, sizeof(st));
    for (int ii = 0; ii < npart; ii++) {
        st.numdata += parts[ii].numdata;
        st.data_avg += parts[ii].data_avg;
        st.data_var += parts[ii].data_var;
        st.prof_avg += parts[ii].prof_avg;
        st.prof_var += parts[ii].prof_var;
    }
    st.data_avg /= npart;
    st.data_var /= npart;
    st.numprof = proflen;
    int bip = (nt - 1) / 2, bipd = (nt - 1) / 2, bipdd = searchpdd ? (nt - 1) / 2 : 0, rc = -1;
    void *stream = NULL, *d_prof = NULL, *d_fd = NULL, *d_pt = NULL, *d_pp = NULL;
    void *d_fdd = NULL;
    if (!nosearch) {
        const size_t pbytes = sizeof(double) * (size_t) npart * proflen;
        if (agpu_init(device) || agpu_stream_create(&stream) || agpu_malloc(&d_prof, pbytes) ||
            agpu_malloc(&d_fd, sizeof(double) * nt) ||
            agpu_malloc(&d_pt, sizeof(double) * npart) ||
            agpu_malloc(&d_fdd, sizeof(double) * nt) ||
            agpu_malloc(&d_pp, sizeof(double) * ncube) ||
            agpu_h2d(d_prof, rawfolds, pbytes, stream) ||
            agpu_h2d(d_fd, fdots, sizeof(double) * nt, stream) ||
            agpu_h2d(d_fdd, fdotdots, sizeof(double) * nt, stream) ||
            agpu_h2d(d_pt, parttimes, sizeof(double) * npart, stream)) {
            accel_set_error("fold search: %s", agpu_last_error());
            goto done;
        }
        agpu_fold_search_args a;
        a.profs = (const double *) d_prof;
        a.fdots = (const double *) d_fd;
        a.fdotdots = searchpdd ? (const double *) d_fdd : NULL;
        a.parttimes = (const double *) d_pt;
        a.prof_avg = st.prof_avg;
        a.prof_var = st.prof_var;
        a.npart = npart;
        a.proflen = proflen;
        a.numtrials = nt;
        a.pstep = pstep;
        for (int ipdd = 0; ipdd < npdd; ipdd++) {
            a.ipdd = ipdd;
            a.redchi = (double *) d_pp + (size_t) ipdd * nt * nt;
            if (agpu_fold_search(&a, stream)) {
                accel_set_error("fold search: %s", agpu_last_error());
                goto done;
            }
        }
        if (agpu_d2h(pp, d_pp, sizeof(double) * ncube, stream) || agpu_stream_sync(stream)) {
            accel_set_error("fold search: %s", agpu_last_error());
            goto done;
        }
        /* the first maximum in (p-dotdot, p-dot, period) order: the
         * reference's strictly-greater test over its loops, from a redchi
         * of 0 */
        double bestrc = 0.0;
        bip = bipd = bipdd = 0;
        for (size_t i = 0; i < ncube; i++)
            if (pp[i] > bestrc) {
                bestrc = pp[i];
                bip = (int) (i % nt);
                bipd = (int) (i / nt % nt);
                bipdd = (int) (i / ((size_t) nt * nt));
            }
        if (ppdot)   /* float, as the reference keeps its plane */
            for (size_t i = 0; i < ncube; i++)
                ppdot[i] = (float) pp[i];
    }
    /* the best profile and its statistics (the redchi in FP64, as
     * combine_profs leaves it) */
    combine_trial(rawfolds, npart, proflen, parttimes, searchpdd ? fdotdots[bipdd] : 0.0,
                  searchpdd, fdots[bipd], pstep * (bip - (nt - 1) / 2), prof);
    double chi = 0.0;
    for (int k = 0; k < proflen; k++) {
        const double d = prof[k] - st.prof_avg;
        chi += d * d;
    }
    st.redchi = chi / st.prof_var / (proflen - 1.0);
    if (bestprof)
        memcpy(bestprof, prof, sizeof(double) * proflen);
    best->numtrials = nt;
    best->ip = bip;
    best->ipd = bipd;
    best->p = periods[bip];
    best->pd = pfdot_of(f, fd + fdots[bipd]);
    best->ipdd = bipdd;
    best->pdd = pfdotdot_of(1.0 / periods[bip], fd + fdots[bipd],
                            searchpdd ? fdd + fdotdots[bipdd] : fdd);
    best->stats = st;
    rc = 0;
done:
    agpu_free(d_prof);
    agpu_free(d_fd);
    agpu_free(d_pt);
    agpu_free(d_pp);
    agpu_free(d_fdd);
    if (stream)
        agpu_stream_destroy(stream);
    free(periods);
    free(fdots);
    free(fdotdots);
    free(pp);
    free(prof);
    return rc;
}
