/*
 * accelsearch_cli.c -- the accelsearch command (src/accelsearch.c:82-313,
 * options of clig/accelsearch_cmd.cli:13-48) over libpresto_accel.
 *
 *   accelsearch [-numharm n] [-zmax z] [-wmax w] [-sigma s] [-flo f] [-fhi f]
 *               [-rlo r] [-rhi r] [-lobin b] [-inmem] [-median] [-photon]
 *               [-zaplist file] [-baryv v] [-noharmremove] [-noharmpolish] [-otheropt]
 *               [-ncpus n] [-device d] file.{dat,sdat,fft}
 *
 * Reads <root>.inf (N, dt, telescope) and the data file; for time series:
 * GPU real FFT, de-reddening and bin 0 := 1+1i (left as de-reddened under
 * -photon), optional birdie zapping; then the search, the FP64 refinement
 * and the outputs <root>_ACCEL_<zmax>[_JERK_<wmax>] (text, .inf appended),
 * .cand and, for an .fft input, .txtcand.  The local-power normalisation
 * (-locpow) scales every bin of a window by its own local power, so the
 * plane of the search handle (raw powers times one norm per window) does
 * not apply: with -locpow the command runs the reference's block loop
 * over the accel.h surface instead (run_surface), which normalises and
 * correlates each window on the GPU.  -otheropt selects the reference's
 * transposed-plane variants, which form the same sums: it changes nothing
 * here.  -ncpus is accepted and ignored (the work runs on the GPU).
 */
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "presto_accel.h"
#include "accel_surface.h"

static void usage(void)
{
    fprintf(stderr,
            "usage: accelsearch [-numharm n] [-zmax z] [-wmax w] [-sigma s] [-flo Hz] [-fhi Hz]\n"
            "                   [-rlo bin] [-rhi bin] [-lobin bin] [-inmem] [-median] [-photon]\n"
            "                   [-zaplist file] [-baryv v] [-noharmremove] [-noharmpolish]\n"
            "                   [-otheropt] [-ncpus n] [-device d] infile.{dat,sdat,fft}\n");
}

/* The option ranges of clig/accelsearch_cmd.cli:13-44, checked as the
 * generated parser does (src/accelsearch_cmd.c:246-335, 640-770). */
static void check_int(const char *opt, long long v, long long lo, long long hi)
{
    if (v > hi) {
        fprintf(stderr, "accelsearch: parameter 1 of option `%s' greater than max=%lld\n", opt, hi);
        exit(EXIT_FAILURE);
    }
    if (v < lo) {
        fprintf(stderr, "accelsearch: parameter 1 of option `%s' smaller than min=%lld\n", opt, lo);
        exit(EXIT_FAILURE);
    }
}

static void check_real(const char *opt, double v, double lo, double hi)
{
    if (v > hi) {
        fprintf(stderr, "accelsearch: parameter 1 of option `%s' greater than max=%f\n", opt, hi);
        exit(EXIT_FAILURE);
    }
    if (v < lo) {
        fprintf(stderr, "accelsearch: parameter 1 of option `%s' smaller than min=%f\n", opt, lo);
        exit(EXIT_FAILURE);
    }
}

typedef struct {
    double N, dt;
    char telescope[100];
} infinfo;

/* The two numbers accelsearch needs from the .inf (src/ioinf.c:294-297),
 * found by their keys. */
static int read_inf(const char *path, infinfo *inf)
{
    FILE *f = fopen(path, "r");
    if (!f)
        return -1;
    char line[512];
    int got = 0;
    inf->telescope[0] = '\0';
    while (fgets(line, sizeof(line), f)) {
        char *eq = strchr(line, '=');
        if (!eq)
            continue;
        const char *val = eq + 1;
        while (*val == ' ')
            val++;
        if (strstr(line, "Number of bins in the time series")) {
            inf->N = atof(val);
            got |= 1;
        } else if (strstr(line, "Width of each time series bin")) {
            inf->dt = atof(val);
            got |= 2;
        } else if (strstr(line, "Telescope used")) {
            snprintf(inf->telescope, sizeof(inf->telescope), "%s", val);
            inf->telescope[strcspn(inf->telescope, "\r\n")] = '\0';
        }
    }
    fclose(f);
    return got == 3 ? 0 : -1;
}

static void *read_file(const char *path, size_t elem, long long *count)
{
    FILE *f = fopen(path, "rb");
    if (!f)
        return NULL;
    fseek(f, 0, SEEK_END);
    long long bytes = ftell(f);
    fseek(f, 0, SEEK_SET);
    *count = bytes / (long long) elem;
    void *buf = malloc((size_t) (*count > 0 ? *count : 1) * elem);
    if (*count && fread(buf, elem, (size_t) *count, f) != (size_t) *count) {
        free(buf);
        fclose(f);
        return NULL;
    }
    fclose(f);
    return buf;
}

/* The reference's driver (src/accelsearch.c:82-290) over the accel.h
 * surface: the path of the -locpow normalisation. */
static int run_surface(Cmdline *cmd)
{
    accelobs obs;
    infodata idata;
    create_accelobs(&obs, &idata, cmd, 1);
    const double rstep = obs.corr_uselen * 0.5;
    /* birdies zapped in the in-memory spectrum (src/accelsearch.c:86-111) */
    if (cmd->zaplistP && !obs.mmap_file && obs.fft) {
        int nb = accel_read_birdies(cmd->zaplist, obs.T, cmd->baryv, NULL, NULL, 0);
        if (nb < 0) {
            fprintf(stderr, "accelsearch: cannot read zaplist '%s'\n", cmd->zaplist);
            return 1;
        }
        double *lo = (double *) malloc(sizeof(double) * (nb ? nb : 1));
        double *hi = (double *) malloc(sizeof(double) * (nb ? nb : 1));
        accel_read_birdies(cmd->zaplist, obs.T, cmd->baryv, lo, hi, nb);
        printf("Zapping them using a barycentric velocity of %.5gc.\n\n", cmd->baryv);
        if (accel_zap_birdies(obs.fft, obs.numbins, (double) obs.N, lo, hi, nb)) {
            fprintf(stderr, "accelsearch: zapping failed\n");
            return 1;
        }
        free(lo);
        free(hi);
    }
    printf("Searching with up to %d harmonics summed:\n", 1 << (obs.numharmstages - 1));
    printf("  f = %.1f to %.1f Hz\n", obs.rlo / obs.T, obs.rhi / obs.T);
    printf("  r = %.1f to %.1f Fourier bins\n", obs.rlo, obs.rhi);
    printf("  z = %.1f to %.1f Fourier bins drifted\n", obs.zlo, obs.zhi);
    if (obs.numw)
        printf("  w = %.1f to %.1f Fourier-derivative bins drifted\n", obs.wlo, obs.whi);
    subharminfo **shis = create_subharminfos(&obs);
    GSList *cands = NULL;
    if (obs.inmem) {
        /* the plane below rlo, for the lowest blocks' sub-harmonics */
        for (double startr = 8.0; startr < obs.rlo; startr += rstep) {
            ffdotpows *f = subharm_fderivs_vol(1, 1, startr, startr + rstep - 0.5, &shis[0][0],
                                               &obs);
            fund_to_ffdotplane(f, &obs);
            free_ffdotpows(f);
        }
    }
    for (double startr = obs.rlo; startr + rstep < obs.highestbin; startr += rstep) {
        const double lastr = startr + rstep - 0.5;
        ffdotpows *fund = subharm_fderivs_vol(1, 1, startr, lastr, &shis[0][0], &obs);
        cands = search_ffdotpows(fund, 1, &obs, cands);
        if (obs.numharmstages > 1) {
            if (obs.inmem)
                fund_to_ffdotplane(fund, &obs);
            for (int stage = 1; stage < obs.numharmstages; stage++) {
                const int h = 1 << stage;
                for (int harm = 1; harm < h; harm += 2) {
                    if (obs.inmem) {
                        inmem_add_ffdotpows(fund, &obs, h, harm);
                    } else {
                        ffdotpows *sub = subharm_fderivs_vol(h, harm, startr, lastr,
                                                             &shis[stage][harm - 1], &obs);
                        add_ffdotpows(fund, sub, h, harm);
                        free_ffdotpows(sub);
                    }
                }
                cands = search_ffdotpows(fund, h, &obs, cands);
            }
        }
        free_ffdotpows(fund);
    }
    free_subharminfos(&obs, shis);
    printf("\nDone searching.  Now optimizing each candidate.\n\n");
    int numcands = 0;
    for (GSList *l = cands; l; l = l->next)
        numcands++;
    if (numcands) {
        cands = sort_accelcands(cands);
        if (cmd->numharm > 1 && !cmd->noharmremoveP)
            cands = eliminate_harmonics(cands, &numcands);
        for (GSList *l = cands; l; l = l->next)
            optimize_accelcand((accelcand *) l->data, &obs);
        fourierprops *props = (fourierprops *) malloc(sizeof(fourierprops) * numcands);
        int i = 0;
        for (GSList *l = cands; l; l = l->next, i++) {
            const accelcand *c = (const accelcand *) l->data;
            /* the fundamental's r and z from the harmonic sum, errors from
             * the grid (src/accelsearch.c:263-275) */
            accel_calc_props(&c->derivs[0], c->r, c->z, c->w, &props[i]);
            props[i].rerr = 0.5f / c->numharm;
            props[i].zerr = 2.0f / c->numharm;
            props[i].werr = 20.0f / c->numharm;
        }
        output_fundamentals(props, cands, &obs, &idata);
        output_harmonics(cands, &obs, &idata);
        FILE *f = fopen(obs.candnm, "wb");
        if (!f || fwrite(props, sizeof(fourierprops), (size_t) numcands, f) != (size_t) numcands) {
            fprintf(stderr, "accelsearch: cannot write '%s'\n", obs.candnm);
            return 1;
        }
        fclose(f);
        free(props);
        printf("Wrote %d candidates to '%s' and '%s'.\n\n", numcands, obs.accelnm, obs.candnm);
    } else {
        printf("No candidates above sigma = %.2f were found.\n\n", obs.sigma);
    }
    for (GSList *l = cands; l; l = l->next)
        free_accelcand(l->data, NULL);
    accel_gslist_free(cands);
    free_accelobs(&obs);
    return 0;
}

int main(int argc, char **argv)
{
    accel_params p;
    accel_params_default(&p);
    const char *infile = NULL, *zaplist = NULL;
    double baryv = 0.0;
    int device = 0, photon = 0, locpow = 0, median = 0, otheropt = 0;
    p.floP = 1;
    p.fhiP = 1;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
#define NEXT() (i + 1 < argc ? argv[++i] : (usage(), exit(1), ""))
        if (!strcmp(a, "-numharm")) {
            p.numharm = atoi(NEXT());
            check_int(a, p.numharm, 1, 32);
        } else if (!strcmp(a, "-zmax")) {
            p.zmax = atoi(NEXT());
            check_int(a, p.zmax, 0, 1200);
        } else if (!strcmp(a, "-wmax")) {
            p.wmax = atoi(NEXT());
            p.wmaxP = 1;
            check_int(a, p.wmax, 0, 4000);
        } else if (!strcmp(a, "-sigma")) {
            p.sigma = (float) atof(NEXT());
            check_real(a, p.sigma, 1.0, 30.0);
        } else if (!strcmp(a, "-flo")) {
            p.flo = atof(NEXT());
            p.floP = 1;
            check_real(a, p.flo, 0.0, HUGE_VAL);
        } else if (!strcmp(a, "-fhi")) {
            p.fhi = atof(NEXT());
            p.fhiP = 1;
            check_real(a, p.fhi, 0.0, HUGE_VAL);
        } else if (!strcmp(a, "-rlo")) {
            /* -flo and -fhi carry defaults, so floP and fhiP are always set
             * (src/accelsearch_cmd.c:49-61) and win over -rlo and -rhi
             * (src/accel_utils.c:1686-1730): those two change nothing */
            p.rlo = atof(NEXT());
            p.rloP = 1;
            check_real(a, p.rlo, 0.0, HUGE_VAL);
        } else if (!strcmp(a, "-rhi")) {
            p.rhi = atof(NEXT());
            p.rhiP = 1;
            check_real(a, p.rhi, 0.0, HUGE_VAL);
        } else if (!strcmp(a, "-lobin")) {
            p.lobin = atoll(NEXT());
            check_int(a, p.lobin, 0, LLONG_MAX);
        } else if (!strcmp(a, "-inmem"))
            p.inmem = 1;
        else if (!strcmp(a, "-median"))
            median = 1;
        else if (!strcmp(a, "-photon"))
            photon = 1;
        else if (!strcmp(a, "-otheropt"))
            otheropt = 1;   /* the transposed-plane variants form the same sums
                 * (src/accel_utils.c:1236-1251, 1284-1309, 1364-1409) */
        else if (!strcmp(a, "-zaplist"))
            zaplist = NEXT();
        else if (!strcmp(a, "-baryv")) {
            baryv = atof(NEXT());
            check_real(a, baryv, -0.1, 0.1);
        }
        else if (!strcmp(a, "-noharmremove"))
            p.noharmremove = 1;
        else if (!strcmp(a, "-ncpus"))
            check_int(a, atoi(NEXT()), 1, INT_MAX);
        else if (!strcmp(a, "-device"))
            device = atoi(NEXT());
        else if (!strcmp(a, "-noharmpolish"))
            p.noharmpolish = 1;
        else if (!strcmp(a, "-locpow"))
            locpow = 1;
        else if (a[0] == '-') {
            fprintf(stderr, "accelsearch: unknown option %s\n", a);
            usage();
            return 1;
        } else
            infile = a;
#undef NEXT
    }
    if (!infile) {
        usage();
        return 1;
    }
    if (locpow) {
        /* the reference's Cmdline, as its generated parser fills it
         * (flo and fhi carry defaults, so they are always given) */
        Cmdline cmd;
        memset(&cmd, 0, sizeof(cmd));
        char *args[1] = {(char *) infile};
        cmd.ncpusP = 1;
        cmd.ncpus = 1;
        cmd.lobinP = p.lobin != 0;
        cmd.lobin = (int) p.lobin;
        cmd.numharmP = 1;
        cmd.numharm = p.numharm;
        cmd.zmaxP = 1;
        cmd.zmax = p.zmax;
        cmd.wmaxP = p.wmaxP;
        cmd.wmax = p.wmax;
        cmd.sigmaP = 1;
        cmd.sigma = p.sigma;
        cmd.rloP = p.rloP;
        cmd.rlo = p.rlo;
        cmd.rhiP = p.rhiP;
        cmd.rhi = p.rhi;
        cmd.floP = 1;
        cmd.flo = p.flo;
        cmd.fhiP = 1;
        cmd.fhi = p.fhi;
        cmd.inmemP = p.inmem;
        cmd.photonP = photon;
        cmd.medianP = median;
        cmd.locpowP = 1;
        cmd.zaplistP = zaplist != NULL;
        cmd.zaplist = (char *) zaplist;
        cmd.baryvP = baryv != 0.0;
        cmd.baryv = baryv;
        cmd.otheroptP = otheropt;
        cmd.noharmpolishP = p.noharmpolish;
        cmd.noharmremoveP = p.noharmremove;
        cmd.argc = 1;
        cmd.argv = args;
        accel_surface_set_device(device);
        return run_surface(&cmd);
    }
    /* root name and input type (src/accel_utils.c:1459-1488) */
    const char *dot = strrchr(infile, '.');
    if (!dot || (strcmp(dot, ".fft") && strcmp(dot, ".dat") && strcmp(dot, ".sdat"))) {
        fprintf(stderr, "\nInput file ('%s') must be an '.fft' or '.[s]dat' file!\n\n", infile);
        return 1;
    }
    char root[4096];
    snprintf(root, sizeof(root), "%.*s", (int) (dot - infile), infile);
    char infpath[4200];
    snprintf(infpath, sizeof(infpath), "%s.inf", root);
    infinfo inf;
    if (read_inf(infpath, &inf)) {
        fprintf(stderr, "accelsearch: cannot read N and dt from '%s'\n", infpath);
        return 1;
    }
    const int dat = strcmp(dot, ".fft") != 0;
    accel_fcomplex *fft = NULL;
    long long numbins = 0;
    clock_t t0 = clock();
    if (dat) {
        long long n = 0;
        float *x = NULL;
        if (!strcmp(dot, ".sdat")) {
            short *sx = (short *) read_file(infile, sizeof(short), &n);
            if (sx) {
                x = (float *) malloc(sizeof(float) * (size_t) (n ? n : 1));
                for (long long k = 0; k < n; k++)
                    x[k] = (float) sx[k];
                free(sx);
            }
        } else {
            x = (float *) read_file(infile, sizeof(float), &n);
        }
        if (!x) {
            fprintf(stderr, "accelsearch: cannot read '%s'\n", infile);
            return 1;
        }
        if (n > 67108864) {   /* src/accel_utils.c:1520 */
            fprintf(stderr, "\nThe input time series is too large.  Use 'realfft' first.\n\n");
            return 1;
        }
        numbins = n / 2;
        fft = (accel_fcomplex *) malloc(sizeof(accel_fcomplex) * (size_t) numbins);
        printf("Reading and FFTing the time series...");
        fflush(stdout);
        if (accel_realfft_forward(x, n, fft, device)) {
            fprintf(stderr, "\naccelsearch: %s\n", accel_last_error());
            return 1;
        }
        free(x);
        printf("done.\nRemoving red-noise...");
        /* src/accel_utils.c:1561 and 1621-1624, the de-reddening on the GPU */
        if (accel_deredden_gpu(fft, numbins, 6, 200, 0.5 / inf.dt, inf.dt * (double) n,
                               device) < 0) {
            fprintf(stderr, "\naccelsearch: de-reddening failed: %s\n", accel_last_error());
            return 1;
        }
        if (!photon) {
            fft[0].r = 1.0f;
            fft[0].i = 1.0f;
        }
        printf("done.\n\n");
        p.N = (double) n;
        p.dat_input = 1;
    } else {
        fft = (accel_fcomplex *) read_file(infile, sizeof(accel_fcomplex), &numbins);
        if (!fft) {
            fprintf(stderr, "accelsearch: cannot read '%s'\n", infile);
            return 1;
        }
        p.N = inf.N;
    }
    p.dt = inf.dt;
    p.numbins = numbins;
    /* src/accel_utils.c:1599-1605: bin 0's real part (mapped .fft, or the
     * de-reddened .dat spectrum) counts the photons */
    if (photon) {
        p.nph = numbins > 0 ? fft[0].r : 0.0;
        printf("Normalizing powers using %.0f photons.\n\n", p.nph);
    }
    if (zaplist && dat) {
        /* src/accelsearch.c:87-110 */
        const double T = p.N * p.dt;
        int nb = accel_read_birdies(zaplist, T, baryv, NULL, NULL, 0);
        if (nb < 0) {
            fprintf(stderr, "accelsearch: cannot read zaplist '%s'\n", zaplist);
            return 1;
        }
        double *lo = (double *) malloc(sizeof(double) * (nb ? nb : 1));
        double *hi = (double *) malloc(sizeof(double) * (nb ? nb : 1));
        accel_read_birdies(zaplist, T, baryv, lo, hi, nb);
        printf("Read %d birdies from '%s'.\n", nb, zaplist);
        printf("Zapping them using a barycentric velocity of %.5gc.\n\n", baryv);
        if (accel_zap_birdies(fft, numbins, p.N, lo, hi, nb)) {
            fprintf(stderr, "accelsearch: zapping failed\n");
            return 1;
        }
        free(lo);
        free(hi);
    }
    accel_search *s = accel_search_create(&p, device);
    if (!s) {
        fprintf(stderr, "accelsearch: %s\n", accel_last_error());
        return 1;
    }
    const accel_geom *g = accel_search_geom(s);
    printf("Searching with up to %d harmonics summed:\n", 1 << (g->numharmstages - 1));
    printf("  f = %.1f to %.1f Hz\n", g->rlo / g->T, g->rhi / g->T);
    printf("  r = %.1f to %.1f Fourier bins\n", g->rlo, g->rhi);
    printf("  z = %.1f to %.1f Fourier bins drifted\n", g->zlo, g->zhi);
    if (g->numw)
        printf("  w = %.1f to %.1f Fourier-derivative bins drifted\n", g->wlo, g->whi);
    char accelnm[4300], candnm[4400], txtnm[4400];
    if (g->numw)
        snprintf(accelnm, sizeof(accelnm), "%s_ACCEL_%d_JERK_%d", root, (int) g->zhi,
                 (int) g->whi);
    else
        snprintf(accelnm, sizeof(accelnm), "%s_ACCEL_%d", root, (int) g->zhi);
    snprintf(candnm, sizeof(candnm), "%s.cand", accelnm);
    snprintf(txtnm, sizeof(txtnm), "%s.txtcand", accelnm);
    /* an .fft input gets the .txtcand work file (src/accel_utils.c:1666-1684) */
    if ((!dat && accel_search_set_txtcand(s, txtnm)) || accel_search_set_spectrum(s, fft) ||
        accel_search_run(s) || accel_search_set_txtcand(s, NULL)) {
        fprintf(stderr, "accelsearch: %s\n", accel_last_error());
        return 1;
    }
    const int ncand = accel_search_ncands(s);
    printf("\nDone searching.  Now optimizing each candidate.\n\n");
    if (ncand) {
        if (accel_search_optimize(s) ||
            accel_search_write_outputs(s, accelnm, candnm, infpath)) {
            fprintf(stderr, "accelsearch: %s\n", accel_last_error());
            return 1;
        }
        printf("Wrote %d candidates to '%s' and '%s'.\n\n", ncand, accelnm, candnm);
    } else {
        printf("No candidates above sigma = %.2f were found.\n\n", g->sigma);
    }
    printf("Searched the following approx numbers of independent points:\n");
    printf("  %d harmonic:   %9lld\n", 1, g->numindep[0]);
    for (int i = 1; i < g->numharmstages; i++)
        printf("  %d harmonics:  %9lld\n", 1 << i, g->numindep[i]);
    float ms[5];
    accel_search_timings(s, ms);
    printf("\nTiming summary:\n    GPU search %.3f ms, total %.2f s (CPU clock)\n", ms[4],
           (double) (clock() - t0) / CLOCKS_PER_SEC);
    accel_search_destroy(s);
    free(fft);
    return 0;
}
