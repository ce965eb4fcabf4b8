/*
 * o_fold.c -- oracle restatement of prepfold's fold of a .dat time series
 * for an ACCEL candidate (prepfold -accelcand, topocentric float data):
 * src/fold.c:46-107 (add_to_prof, the "drizzling" of finite-duration
 * samples into profile bins through a buffer that is dumped once a full
 * turn of phase has been added), src/fold.c:443-653 (fold(), flags 0, in
 * long double as there), src/prepfold_pipeline.c:514-551 and 1375-1594
 * (records of SUBSBLOCKLEN = 1024 points, npart sub-integrations of
 * reads_per_part records, buffers and phases carried across calls), and
 * src/prepfold_pipeline.c:750-767 (the start f, fdot, fdotdot of a
 * candidate's average r, z, w).  Test infrastructure only (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define O_SUBSBLOCKLEN 1024
#define O_TEST_ONE(a) (fabs((a) - 1.0) < 1e-12)

/* src/fold.c:46-107 */
static void add_to_prof(double *prof, double *buffer, int n, double lophase,
                        double deltaphase, double dataval, double *phaseadded)
{
    const double binw = 1.0 / n;
    const double valperphs = dataval / deltaphase;
    double curbin = lophase * n;
    while (deltaphase > 1e-12) {
        int ib = (int) floor(curbin + 1e-12);
        double dphs = ((ib + 1.0) - curbin) * binw;
        ib %= n;
        if (dphs > deltaphase)
            dphs = deltaphase;
        const double room = 1.0 - *phaseadded;
        double part = room > dphs ? dphs : room;
        buffer[ib] += part * valperphs;
        *phaseadded += part;
        if (O_TEST_ONE(*phaseadded)) {
            for (int k = 0; k < n; k++) {
                prof[k] += buffer[k];
                buffer[k] = 0.0;
            }
            part = dphs - room;
            buffer[ib] += part * valperphs;
            *phaseadded = part;
        }
        deltaphase -= dphs;
        curbin += dphs * n;
    }
}

/* src/fold.c:443-653 with flags = 0 (no delays, no on/off pairs) and the
 * standard drizzling; stats accumulate across calls as there */
static double fold_block(const float *data, int numdata, double dt, double tlo, double *prof,
                         int numprof, double startphs, double *buffer, double *phaseadded,
                         double fo, double fdot, double fdotdot, o_foldstats *st)
{
    long double phase, phasenext = 0.0, deltaphase, T, Tnext, lophase, hiphase;
    double dtmp;
    fdot /= 2.0;
    fdotdot /= 6.0;
    st->numprof = (double) numprof;
    st->data_var *= (st->numdata - 1.0);
    T = tlo;
    phase = T * (T * (T * fdotdot + fdot) + fo) + startphs;
    lophase = (phase < 0.0) ? 1.0 + modf((double) phase, &dtmp) : modf((double) phase, &dtmp);
    for (int ii = 0; ii < numdata; ii++) {
        Tnext = tlo + (ii + 1) * dt;
        phasenext = Tnext * (Tnext * (Tnext * fdotdot + fdot) + fo) + startphs;
        deltaphase = phasenext - phase;
        add_to_prof(prof, buffer, numprof, (double) lophase, (double) deltaphase, data[ii],
                    phaseadded);
        hiphase = lophase + deltaphase;
        lophase = hiphase - (int) hiphase;
        phase = phasenext;
        st->numdata += 1.0;
        const double dev = data[ii] - st->data_avg;
        st->data_avg += dev / st->numdata;
        st->data_var += dev * (data[ii] - st->data_avg);
    }
    st->prof_avg = 0.0;
    for (int k = 0; k < numprof; k++)
        st->prof_avg += prof[k];
    st->prof_avg /= numprof;
    st->redchi = 0.0;
    for (int k = 0; k < numprof; k++) {
        dtmp = prof[k] - st->prof_avg;
        st->redchi += dtmp * dtmp;
    }
    st->data_var /= (st->numdata - 1.0);
    st->prof_var = st->data_var * st->numdata / numprof;
    st->redchi /= (st->prof_var * (numprof - 1));
    phasenext = (phasenext < 0.0) ? 1.0 + phasenext - (int) phasenext
                                  : phasenext - (int) phasenext;
    return (double) phasenext;
}

/* src/prepfold_pipeline.c:750-767 */
void o_cand_fold_params(double r, double z, double w, double T, double *f, double *fd,
                        double *fdd)
{
    const double z0 = z - 0.5 * w;
    const double r0 = r - 0.5 * z0 - w / 6.0;
    *f = r0 / T;
    *fd = z0 / (T * T);
    *fdd = w / (T * T * T);
}

/* The fold of fold_timeseries() for a .dat of n floats: the records used
 * (numrec = n / 1024 rounded down to a multiple of npart), every record
 * folded at its own start time, one profile and statistics per part.
 * rawfolds: npart x proflen (zeroed here); stats: npart.  Returns the
 * number of points folded. */
long long o_fold_dat(const float *series, long long n, double dt, double f, double fd,
                     double fdd, int npart, int proflen, double *rawfolds, o_foldstats *stats)
{
    const long long numrec0 = n / O_SUBSBLOCKLEN;
    const long long rpp = numrec0 / npart;
    double *buffer = (double *) calloc(proflen, sizeof(double));
    double phaseadded = 0.0;
    memset(rawfolds, 0, sizeof(double) * (size_t) npart * proflen);
    memset(stats, 0, sizeof(o_foldstats) * (size_t) npart);
    const double proftime = O_SUBSBLOCKLEN * dt;
    long long done = 0;
    for (int ii = 0; ii < npart; ii++) {
        const double parttime = ii * rpp * proftime;
        for (long long jj = 0; jj < rpp; jj++) {
            const double t0 = parttime + jj * proftime;
            fold_block(series + done, O_SUBSBLOCKLEN, dt, t0, rawfolds + (size_t) ii * proflen,
                       proflen, 0.0, buffer, &phaseadded, f, fd, fdd, &stats[ii]);
            done += O_SUBSBLOCKLEN;
        }
    }
    free(buffer);
    return done;
}

/* ---- prepfold's period / period-derivative search, one subband ----
 * src/prepfold_pipeline.c:1633-1915 with nsub 1, no p-dotdot search;
 * src/fold.c:696-754 (combine_profs); src/characteristics.c:562-575
 * (chisqr); src/prepfold_utils.c:7-73 (the phase-delay conversions). */

static double o_fdot2phasedelay(double fdot, double time)
{
    double r = fdot * time * time / 2.0;
    return r == -0 ? 0.0 : r;
}

static double o_fdotdot2phasedelay(double fdotdot, double time)
{
    double r = fdotdot * time * time * time / 6.0;
    return r == -0 ? 0.0 : r;
}

static double o_phasedelay2fdotdot(double phasedelay, double time)
{
    return time == 0.0 ? 0.0 : 6.0 * phasedelay / (time * time * time);
}

static double o_phasedelay2fdot(double phasedelay, double time)
{
    return time == 0.0 ? 0.0 : 2.0 * phasedelay / (time * time);
}

static double o_switch_pfdot(double pf, double pfdot)
{
    if (pf == 0.0)
        return 0.0;
    double r = -pfdot / (pf * pf);
    return r == -0 ? 0.0 : r;
}

static double o_switch_pfdotdot(double pf, double pfdot, double pfdotdot)
{
    if (pf == 0.0 || pfdotdot == 0.0)
        return 0.0;
    double r = 2.0 * pfdot * pfdot / (pf * pf * pf) - pfdotdot / (pf * pf);
    return r == -0 ? 0.0 : r;
}

/* combine_profs: each profile shifted by its delay rounded to whole bins
 * and summed in part order; the parts' profile statistics summed; the
 * reduced chi-squared of the sum */
static double o_combine(const double *profs, int npart, int proflen, const double *delays,
                        double prof_avg, double prof_var, double *out)
{
    for (int k = 0; k < proflen; k++)
        out[k] = 0.0;
    int index = 0;
    for (int ii = 0; ii < npart; ii++) {
        double ld = fmod(-delays[ii], proflen);
        if (ld < 0.0)
            ld += proflen;
        const int offset = (int) (ld + 0.5);
        int jj = 0, kk = proflen - offset;
        for (; jj < offset; jj++, kk++, index++)
            out[kk] += profs[index];
        for (kk = 0; jj < proflen; jj++, kk++, index++)
            out[kk] += profs[index];
    }
    double chi = 0.0;
    for (int k = 0; k < proflen; k++) {
        const double d = out[k] - prof_avg;
        chi += d * d;
    }
    return chi / prof_var / (proflen - 1.0);
}

int o_fold_search(const double *profs, const o_foldstats *parts, int npart, int proflen,
                  double f, double fd, double fdd, double T, const double *parttimes,
                  int npfact, int pstep, int pdstep, int searchpdd, float *ppdot, double *best)
{
    const int nt = 2 * npfact * proflen + 1;
    double prof_avg = 0.0, prof_var = 0.0;
    for (int ii = 0; ii < npart; ii++) {
        prof_avg += parts[ii].prof_avg;
        prof_var += parts[ii].prof_var;
    }
    double *fdots = (double *) malloc(sizeof(double) * nt);
    double *periods = (double *) malloc(sizeof(double) * nt);
    double *pdots = (double *) malloc(sizeof(double) * nt);
    double *fdotdots = (double *) malloc(sizeof(double) * nt);
    double *pdd_delays = (double *) malloc(sizeof(double) * npart);
    double *pd_delays = (double *) malloc(sizeof(double) * npart);
    double *delays = (double *) malloc(sizeof(double) * npart);
    double *cur = (double *) malloc(sizeof(double) * proflen);
    for (int i = 0; i < nt; i++) {
        const int tot = i - (nt - 1) / 2;
        double d = (double) (tot * pstep) / proflen;
        periods[i] = 1.0 / (f + d / T);
        d = (double) (tot * pdstep) / proflen;
        fdots[i] = o_phasedelay2fdot(d, T);
        pdots[i] = o_switch_pfdot(f, fd + fdots[i]);
        fdotdots[i] = o_phasedelay2fdotdot(d, T);
    }
    /* src/prepfold_pipeline.c:1814-1868 for one subband: the p-dotdot
     * loop outermost when searched, one pass with no p-dotdot delay
     * otherwise; the first strictly larger redchi wins */
    const int npdd = searchpdd ? nt : 1;
    int bip = 0, bipd = 0, bipdd = 0;
    double bestredchi = 0.0;
    for (int ipdd = 0; ipdd < npdd; ipdd++) {
        for (int ii = 0; ii < npart; ii++)
            pdd_delays[ii] =
                searchpdd ? o_fdotdot2phasedelay(fdotdots[ipdd], parttimes[ii]) * proflen : 0.0;
        for (int ipd = 0; ipd < nt; ipd++) {
            for (int ii = 0; ii < npart; ii++)
                pd_delays[ii] =
                    pdd_delays[ii] + o_fdot2phasedelay(fdots[ipd], parttimes[ii]) * proflen;
            for (int ip = 0; ip < nt; ip++) {
                const int totpdelay = pstep * (ip - (nt - 1) / 2);
                for (int ii = 0; ii < npart; ii++)
                    delays[ii] = pd_delays[ii] + (double) (ii * totpdelay) / npart;
                const double rc =
                    o_combine(profs, npart, proflen, delays, prof_avg, prof_var, cur);
                if (ppdot)
                    ppdot[((size_t) ipdd * nt + ipd) * nt + ip] = (float) rc;
                if (rc > bestredchi) {
                    bestredchi = rc;
                    bip = ip;
                    bipd = ipd;
                    bipdd = ipdd;
                }
            }
        }
    }
    /* best[]: ip, ipd, redchi, p, pd, pdd, ipdd (src/prepfold_pipeline.c:1886-1897) */
    best[0] = bip;
    best[1] = bipd;
    best[2] = bestredchi;
    best[3] = periods[bip];
    best[4] = pdots[bipd];
    best[5] = o_switch_pfdotdot(1.0 / periods[bip], fd + fdots[bipd],
                                searchpdd ? fdd + fdotdots[bipdd] : fdd);
    best[6] = bipdd;
    free(fdots);
    free(periods);
    free(pdots);
    free(pd_delays);
    free(fdotdots);
    free(pdd_delays);
    free(delays);
    free(cur);
    return nt;
}
