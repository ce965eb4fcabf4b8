/*
 * accelfold -- prepfold -accelcand for a .dat time series, on the GPU:
 *
 *     accelfold [-n proflen] [-npart n] [-npfact n] [-pstep n] [-pdstep n]
 *               [-nosearch] [-searchpdd] -accelfile X_ACCEL_200.cand -accelcand k X.dat
 *
 * Reads record k of the .cand file (src/get_candidates.c:27-37) and the
 * .inf of its root (the part before "_ACCEL", src/prepfold_pipeline.c:
 * 745-756), turns the candidate's average r, z, w into the start f, fdot,
 * fdotdot (src/prepfold_pipeline.c:757-767), folds the series as prepfold
 * does (accel_fold) and, unless -nosearch, searches the sub-integrations
 * for the best period and p-dot, and with -searchpdd p-dotdot
 * (accel_fold_search,
 * src/prepfold_pipeline.c:1633-1915 for one subband).  Prints the folding
 * parameters, the best trial, its profile and reduced chi-squared.
 * prepfold's option names and defaults are kept
 * (clig/prepfold_cmd.cli:45,55-67); its DM search, plots and .pfd output
 * are not built, and only the options above are accepted.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "presto_accel.h"
#include "accel_surface.h"

static void usage(void)
{
    fprintf(stderr, "usage: accelfold [-n proflen] [-npart n] [-npfact n] [-pstep n] "
                    "[-pdstep n] [-nosearch] [-searchpdd] -accelfile file.cand -accelcand k file.dat\n");
    exit(1);
}

int main(int argc, char **argv)
{
    const char *candfile = NULL, *datfile = NULL;
    int candnum = 0, proflen = 0, npart = 64, npfact = 2, pstep = 2, pdstep = 4, nosearch = 0;
    int searchpdd = 0;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "-accelfile") && i + 1 < argc)
            candfile = argv[++i];
        else if (!strcmp(argv[i], "-accelcand") && i + 1 < argc)
            candnum = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-n") && i + 1 < argc)
            proflen = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-npart") && i + 1 < argc)
            npart = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-npfact") && i + 1 < argc)
            npfact = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-pstep") && i + 1 < argc)
            pstep = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-pdstep") && i + 1 < argc)
            pdstep = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-nosearch"))
            nosearch = 1;
        else if (!strcmp(argv[i], "-searchpdd"))
            searchpdd = 1;
        else if (argv[i][0] == '-')
            usage();
        else
            datfile = argv[i];
    }
    if (!candfile || !datfile || candnum < 1)
        usage();
    if (npart < 1 || npart > 4096 || (proflen && (proflen < 2 || proflen > 1 << 20))) {
        fprintf(stderr, "accelfold: -npart must be 1..4096 and -n 2..1048576\n");
        return 1;
    }
    if (npfact < 1 || npfact > 10 || pstep < 1 || pstep > 10 || pdstep < 1 || pdstep > 20) {
        /* the reference's ranges, clig/prepfold_cmd.cli:60-67 */
        fprintf(stderr, "accelfold: -npfact must be 1..10, -pstep 1..10 and -pdstep 1..20\n");
        return 1;
    }
    /* the .inf of the candidate file's root */
    char root[4096];
    snprintf(root, sizeof(root), "%s", candfile);
    char *cut = strstr(root, "_ACCEL");
    if (!cut) {
        fprintf(stderr, "accelfold: '%s' is not an ACCEL candidate file\n", candfile);
        return 1;
    }
    *cut = '\0';
    infodata idata;
    if (accel_readinf(&idata, root)) {
        fprintf(stderr, "accelfold: %s\n", accel_last_error());
        return 1;
    }
    accel_fourierprops cand;
    if (accel_read_rzw_cand(candfile, candnum, &cand)) {
        fprintf(stderr, "accelfold: %s\n", accel_last_error());
        return 1;
    }
    const double T = idata.dt * idata.N;
    double f, fd, fdd;
    accel_cand_fold_params(&cand, T, &f, &fd, &fdd);
    if (!proflen)
        proflen = accel_fold_proflen(f, idata.dt);
    FILE *fp = fopen(datfile, "rb");
    if (!fp) {
        fprintf(stderr, "accelfold: cannot open '%s'\n", datfile);
        return 1;
    }
    fseek(fp, 0, SEEK_END);
    const long long n = ftell(fp) / (long long) sizeof(float);
    fseek(fp, 0, SEEK_SET);
    float *x = (float *) malloc(sizeof(float) * (size_t) (n > 0 ? n : 1));
    if (fread(x, sizeof(float), (size_t) n, fp) != (size_t) n) {
        fprintf(stderr, "accelfold: short read of '%s'\n", datfile);
        return 1;
    }
    fclose(fp);
    double *prof = (double *) malloc(sizeof(double) * (size_t) npart * proflen);
    double *best_prof = (double *) malloc(sizeof(double) * proflen);
    double *parttimes = (double *) malloc(sizeof(double) * npart);
    accel_foldstats *parts = (accel_foldstats *) malloc(sizeof(accel_foldstats) * npart);
    accel_foldstats total;
    const long long used =
        accel_fold(x, n, idata.dt, f, fd, fdd, npart, proflen, prof, parts, &total, 0);
    if (used < 0) {
        fprintf(stderr, "accelfold: %s\n", accel_last_error());
        return 1;
    }
    /* sub-integrations start at whole reads of 1024 points, and the span
     * is what was folded (src/prepfold_pipeline.c:545-552,1515) */
    const long long rpp = used / 1024 / npart;
    for (int p = 0; p < npart; p++)
        parttimes[p] = (double) (p * rpp) * 1024.0 * idata.dt;
    accel_fold_best best;
    if (accel_fold_search(prof, parts, npart, proflen, f, fd, fdd, used * idata.dt, parttimes,
                          npfact, pstep, pdstep, nosearch, searchpdd, NULL, best_prof, &best,
                          0)) {
        fprintf(stderr, "accelfold: %s\n", accel_last_error());
        return 1;
    }
    printf("Candidate %d of '%s': r = %.6f, z = %.4f, w = %.4f (sigma %.2f)\n", candnum,
           candfile, cand.r, cand.z, cand.w, cand.sig);
    printf("Folding %lld of %lld points (%d parts, %d bins)\n", used, n, npart, proflen);
    printf("  f = %.12g Hz   fd = %.6g Hz/s   fdd = %.6g Hz/s^2\n", f, fd, fdd);
    printf("  p = %.12g s    pd = %.6g s/s\n", 1.0 / f, -fd / (f * f));
    printf("Fold reduced chi-squared = %.4f\n", total.redchi);
    if (!nosearch)
        printf("Searched %d periods x %d p-dots%s (npfact %d, pstep %d, pdstep %d)\n",
               best.numtrials, best.numtrials, searchpdd ? " x p-dotdots" : "", npfact, pstep,
               pdstep);
    printf("Best p = %.15g s   pd = %.8g s/s   pdd = %.8g s/s^2  (trial %d, %d, %d)\n", best.p,
           best.pd, best.pdd, best.ip, best.ipd, best.ipdd);
    printf("Profile (%s):\n", nosearch ? "sum of the parts" : "best trial");
    for (int k = 0; k < proflen; k++)
        printf("  %4d  %.6f\n", k, best_prof[k]);
    printf("Data: avg %.6g, var %.6g\n", best.stats.data_avg, best.stats.data_var);
    printf("Reduced chi-squared = %.4f\n", best.stats.redchi);
    free(x);
    free(prof);
    free(best_prof);
    free(parttimes);
    free(parts);
    return 0;
}
