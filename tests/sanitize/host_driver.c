/*
 * Host-only driver for the sanitizer build of the library's C host code
 * (geometry, statistics, de-reddening, birdie zapping, the ACCEL output
 * formatting, the sifter and the .cand reader), built by
 * tests/test_sanitize.py with -fsanitize=address,undefined.  No device
 * code is linked: only host entry points are called.  argv[1]: a directory
 * of ACCEL files (the sifting fixtures); argv[2]: a scratch directory.
 * Exits 0 when every check holds; the sanitizers abort on the first error.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <dirent.h>
#include "presto_accel.h"

#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            return 1;                                              \
        }                                                          \
    } while (0)

static int geometry(void)
{
    /* the configurations' shapes and some odd ones */
    const double Ns[] = {1 << 20, 1 << 23, 1 << 22, 1 << 16, 3000000, 1 << 24};
    const int zs[] = {50, 200, 200, 0, 51, 200}, ws[] = {0, 0, 300, 0, 0, 100};
    const int nhs[] = {8, 16, 8, 1, 32, 8};
    for (int k = 0; k < 6; k++) {
        accel_params p;
        accel_geom g;
        accel_params_default(&p);
        p.N = Ns[k];
        p.dt = 64e-6;
        p.numbins = (long long) Ns[k] / 2;
        p.zmax = zs[k];
        p.numharm = nhs[k];
        if (ws[k]) {
            p.wmax = ws[k];
            p.wmaxP = 1;
        }
        p.sigma = 2.0f;
        CHECK(accel_geom_create(&p, &g) == 0);
        CHECK(g.nblocks > 0 && g.corr_uselen > 0 && g.fftlen >= g.corr_uselen);
    }
    return 0;
}

static int stats(void)
{
    /* powers above the search thresholds: sigma -> power -> sigma */
    for (int ns = 1; ns <= 32; ns *= 2)
        for (double sig = 2.0; sig < 30.0; sig += 1.5) {
            const double pw = accel_power_for_sigma(sig, ns, 1e9);
            CHECK(isfinite(pw) && pw > 0.0);
            const double back = accel_candidate_sigma(pw, ns, 1e9);
            CHECK(fabs(back - sig) < 1e-3 * sig);
        }
    return 0;
}

static int dered_and_zap(void)
{
    const long n = 1 << 16;
    accel_fcomplex *f = (accel_fcomplex *) malloc(sizeof(accel_fcomplex) * n);
    unsigned int u[4];
    CHECK(accel_mt19937_raw(1, u, 4) == 4);
    for (long i = 0; i < n; i++) {
        f[i].r = (float) (((i * 7919) % 101) - 50) / 25.0f;
        f[i].i = (float) (((i * 104729) % 97) - 48) / 24.0f;
    }
    CHECK(accel_prepare_dat_spectrum(f, n, 2.0 * n, 1e-3) > 0);
    double lo[2] = {100.0, 2000.5}, hi[2] = {103.0, 2010.0};
    CHECK(accel_zap_birdies(f, n, 2.0 * n, lo, hi, 2) >= 0);
    for (long i = 0; i < n; i++)
        CHECK(isfinite(f[i].r) && isfinite(f[i].i));
    free(f);
    return 0;
}

static int output(void)
{
    char buf[96];
    const double vals[] = {0.0, 1.5, -123.456, 1e-9, 7.25e7, 99999.9};
    for (int i = 0; i < 6; i++)
        for (int ne = 1; ne <= 2; ne++) {
            int rc = accel_nice_output(buf, vals[i], fabs(vals[i]) * 0.01 + 1e-3, 0, ne);
            CHECK(rc >= 0 && strlen(buf) < sizeof(buf));
        }
    accel_rderivs d = {50.0, 1.0, -0.5, 0.2, -30.0, 0.01, 1.0};
    accel_fourierprops fp;
    accel_rzwerrs e;
    CHECK(accel_calc_props(&d, 1000.25, 3.0, 0.0, &fp) == 0);
    CHECK(accel_calc_rzwerrs(&fp, 100.0, &e) == 0);
    return 0;
}

static int sift(const char *dir, const char *tmp)
{
    accel_sift *s = accel_sift_create(1, 1);
    CHECK(s);
    char path[1024];
    int total = 0, files = 0;
    DIR *d = opendir(dir);
    CHECK(d);
    struct dirent *e;
    while ((e = readdir(d))) {
        if (!strstr(e->d_name, "_ACCEL_") || strstr(e->d_name, ".cand"))
            continue;
        snprintf(path, sizeof(path), "%s/%s", dir, e->d_name);
        int n = accel_sift_read(s, path);
        if (n < 0)
            fprintf(stderr, "cannot read %s\n", path);
        CHECK(n >= 0);
        total += n;
        files++;
    }
    closedir(d);
    CHECK(files > 0 && total > 0);
    accel_sift_params p;
    memset(&p, 0, sizeof(p));
    p.r_err = 1.1;
    p.long_period = 15.0;
    p.short_period = 0.0005;
    p.sigma_threshold = 2.0;
    p.c_pow_threshold = 100.0;
    p.harm_pow_cutoff = 8.0;
    CHECK(accel_sift_reject(s, &p, ACCEL_SIFT_R_DEFAULT) >= 0);
    CHECK(accel_sift_remove_duplicates(s, &p) >= 0);
    const double dms[] = {0, 10, 20, 30, 40, 50};
    CHECK(accel_sift_remove_dm_problems(s, 2, dms, 6, 2.0) >= 0);
    CHECK(accel_sift_remove_harmonics(s, &p) >= 0);
    accel_sift_sort_sigma(s);
    snprintf(path, sizeof(path), "%s/sifted.txt", tmp);
    CHECK(accel_sift_write(s, path) == 0);
    snprintf(path, sizeof(path), "%s/report.txt", tmp);
    CHECK(accel_sift_write_report(s, path) == 0);
    snprintf(path, sizeof(path), "%s/summary.txt", tmp);
    CHECK(accel_sift_write_summary(s, path) == 0);
    CHECK(accel_sift_parse_power("12.5") > 12.0);
    accel_sift_destroy(s);
    return 0;
}

static int cand_reader(const char *tmp)
{
    accel_fourierprops recs[3];
    memset(recs, 0, sizeof(recs));
    for (int k = 0; k < 3; k++)
        recs[k].r = 100.0 + k;
    char path[1024];
    snprintf(path, sizeof(path), "%s/x_ACCEL_0.cand", tmp);
    FILE *f = fopen(path, "wb");
    CHECK(f && fwrite(recs, sizeof(recs[0]), 3, f) == 3);
    fclose(f);
    accel_fourierprops c;
    CHECK(accel_read_rzw_cand(path, 3, &c) == 0 && c.r == 102.0);
    CHECK(accel_read_rzw_cand(path, 4, &c) == -1);
    double fr, fd, fdd;
    accel_cand_fold_params(&c, 10.0, &fr, &fd, &fdd);
    CHECK(fr > 0.0);
    CHECK(accel_fold_proflen(fr, 1e-3) >= 2);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc < 3)
        return 2;
    if (geometry() || stats() || dered_and_zap() || output() || sift(argv[1], argv[2]) ||
        cand_reader(argv[2]))
        return 1;
    printf("sanitized host checks passed\n");
    return 0;
}
