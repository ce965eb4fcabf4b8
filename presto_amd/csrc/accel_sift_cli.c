/*
 * accel_sift_cli.c -- the accel_sift command: the sifting flow of PRESTO's
 * examplescripts/ACCEL_sift.py over the ACCEL files of a directory, on the
 * C sifter of libpresto_accel.so.
 *
 *   accel_sift [-o file] [-glob pattern] [-infglob pattern] [-sigma s]
 *              [-cpow p] [-min_num_DMs n] [-low_DM_cutoff dm] [-r_err r]
 *              [-short_period s] [-long_period s] [-harm_pow_cutoff p]
 *              [-report file] [-summary]
 *
 * The DM list comes from the .inf file names (…DM<dm>.inf), or from the
 * ACCEL file names when the first .inf name marks a short search
 * ("_<3 digits>M_"), as ACCEL_sift.py takes it.  Every ACCEL file is read
 * and given the default rejection; then duplicates across DMs are merged,
 * candidates with DM problems and harmonics of stronger candidates are set
 * aside, and the rest is written in sigma order.  Files are taken in name
 * order (the reference takes them in directory order, which makes ties
 * between equal frequencies depend on the file system).
 */
#include <glob.h>
#include <regex.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "presto_accel.h"

static void usage(void)
{
    fprintf(stderr,
            "usage: accel_sift [-o file] [-glob '*ACCEL_*0'] [-infglob '*DM*.inf'] [-sigma 4.0]\n"
            "                  [-cpow 100] [-min_num_DMs 2] [-low_DM_cutoff 2.0] [-r_err 1.1]\n"
            "                  [-short_period 0.0005] [-long_period 15.0] [-harm_pow_cutoff 8.0]\n"
            "                  [-report file] [-summary]\n");
}

static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *) a, y = *(const double *) b;
    return (x > y) - (x < y);
}

/* the text between the last "DM" of s and the first `stop` after it */
static double dm_after_last(const char *s, const char *stop)
{
    const char *p = NULL;
    for (const char *q = strstr(s, "DM"); q; q = strstr(q + 1, "DM"))
        p = q;
    if (!p)
        return -1.0;
    char buf[64];
    snprintf(buf, sizeof(buf), "%s", p + 2);
    char *e = strstr(buf, stop);
    if (e)
        *e = '\0';
    return strtod(buf, NULL);
}

int main(int argc, char **argv)
{
    const char *out = NULL, *report = NULL, *gaccel = "*ACCEL_*0", *ginf = "*DM*.inf";
    int min_num_dms = 2, summary = 0;
    double low_dm_cutoff = 2.0;
    accel_sift_params p = {1.1, 15.0, 0.0005, 4.0, 100.0, 8.0, NULL, 0, NULL, 0};
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        const int more = i + 1 < argc;
        if (!strcmp(a, "-o") && more)
            out = argv[++i];
        else if (!strcmp(a, "-report") && more)
            report = argv[++i];
        else if (!strcmp(a, "-glob") && more)
            gaccel = argv[++i];
        else if (!strcmp(a, "-infglob") && more)
            ginf = argv[++i];
        else if (!strcmp(a, "-sigma") && more)
            p.sigma_threshold = atof(argv[++i]);
        else if (!strcmp(a, "-cpow") && more)
            p.c_pow_threshold = atof(argv[++i]);
        else if (!strcmp(a, "-min_num_DMs") && more)
            min_num_dms = atoi(argv[++i]);
        else if (!strcmp(a, "-low_DM_cutoff") && more)
            low_dm_cutoff = atof(argv[++i]);
        else if (!strcmp(a, "-r_err") && more)
            p.r_err = atof(argv[++i]);
        else if (!strcmp(a, "-short_period") && more)
            p.short_period = atof(argv[++i]);
        else if (!strcmp(a, "-long_period") && more)
            p.long_period = atof(argv[++i]);
        else if (!strcmp(a, "-harm_pow_cutoff") && more)
            p.harm_pow_cutoff = atof(argv[++i]);
        else if (!strcmp(a, "-summary"))
            summary = 1;
        else {
            usage();
            return 1;
        }
    }
    glob_t ga, gi;
    memset(&ga, 0, sizeof(ga));
    memset(&gi, 0, sizeof(gi));
    glob(gaccel, 0, NULL, &ga);   /* sorted by name */
    glob(ginf, 0, NULL, &gi);
    if (gi.gl_pathc == 0) {
        fprintf(stderr, "accel_sift: no .inf files match %s\n", ginf);
        globfree(&ga);
        globfree(&gi);
        return 1;
    }
    /* the DM list */
    regex_t shortsearch;
    regcomp(&shortsearch, "_[0-9][0-9][0-9]M_", REG_NOSUB);
    const int from_accel = regexec(&shortsearch, gi.gl_pathv[0], 0, NULL, 0) == 0;
    regfree(&shortsearch);
    const size_t ndm = from_accel ? ga.gl_pathc : gi.gl_pathc;
    double *dms = (double *) malloc(sizeof(double) * (ndm ? ndm : 1));
    for (size_t i = 0; i < ndm; i++)
        dms[i] = from_accel ? dm_after_last(ga.gl_pathv[i], "_")
                            : dm_after_last(gi.gl_pathv[i], ".inf");
    qsort(dms, ndm, sizeof(double), cmp_double);
    int rc = 0;
    accel_sift *all = accel_sift_create(report != NULL || summary, report != NULL || summary);
    for (size_t i = 0; i < ga.gl_pathc && rc == 0; i++) {
        accel_sift *one = accel_sift_create(report != NULL || summary, report != NULL || summary);
        if (accel_sift_read(one, ga.gl_pathv[i]) < 0 ||
            accel_sift_reject(one, &p, ACCEL_SIFT_R_DEFAULT) < 0 || accel_sift_extend(all, one))
            rc = 1;
        accel_sift_destroy(one);
    }
    if (rc == 0 && accel_sift_count(all, ACCEL_SIFT_GOOD) > 0)
        rc = accel_sift_remove_duplicates(all, &p) < 0;
    if (rc == 0 && accel_sift_count(all, ACCEL_SIFT_GOOD) > 0)
        rc = accel_sift_remove_dm_problems(all, min_num_dms, dms, (int) ndm, low_dm_cutoff) < 0;
    if (rc == 0 && accel_sift_count(all, ACCEL_SIFT_GOOD) > 0)
        rc = accel_sift_remove_harmonics(all, &p) < 0;
    if (rc == 0 && accel_sift_count(all, ACCEL_SIFT_GOOD) > 0) {
        accel_sift_sort_sigma(all);
        rc = accel_sift_write(all, out) < 0;
    }
    if (rc == 0 && report)
        rc = accel_sift_write_report(all, report) < 0;
    if (rc == 0 && summary)
        rc = accel_sift_write_summary(all, NULL) < 0;
    if (rc)
        fprintf(stderr, "accel_sift: %s\n", accel_last_error());
    accel_sift_destroy(all);
    free(dms);
    globfree(&ga);
    globfree(&gi);
    return rc;
}
