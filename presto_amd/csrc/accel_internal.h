/*
 * accel_internal.h -- functions shared between the host sources of
 * libpresto_accel.so (accel_host.c, accel_output.c, accel_surface.c).
 * Not part of the public ABI: the symbols are hidden.
 */
#ifndef PRESTO_ACCEL_INTERNAL_H
#define PRESTO_ACCEL_INTERNAL_H

#include <stdio.h>
#include "presto_accel.h"

#define ACCEL_HIDDEN __attribute__((visibility("hidden")))

#ifdef __cplusplus
extern "C" {
#endif

/* Device state of a search handle that the accel.h surface works on. */
typedef struct {
    const void *d_spec;   /* numbins complex FP32, bin lobin first */
    long long numbins, lobin;
    const float *d_plane; /* raw plane, AGPU_PLANE_IDX layout */
    long long ntiles, hb_end;
    int nrows, numz, numw1;
    void *stream;
} accel_dev_view;

ACCEL_HIDDEN void accel_search_dev_view(const accel_search *s, accel_dev_view *v);
/* accel_search_create with nstreams = 2 (the block norms on a second
 * stream beside the correlation) or 1 (everything on one stream: for the
 * request loop, whose concurrent requests already overlap, and where more
 * streams per handle would share the device's few hardware queues). */
ACCEL_HIDDEN accel_search *accel_search_create_streams(const accel_params *p, int device,
                                                       int nstreams);
/* The raw plane alone (correlation and halos over every half-bin), then
 * waits for it. */
ACCEL_HIDDEN int accel_search_plane_only(accel_search *s);
/* Harmonic-polished refinement of n candidates on the handle's resident
 * spectrum (rin relative to lobin; win NULL for rz); pows and derivs are
 * [n][ACCEL_MAXHARM].  raw: single harmonics without polishing
 * (max_rz_arr / max_rzw_arr; every nh must be 1). */
ACCEL_HIDDEN int accel_search_refine_dev(accel_search *s, int raw, int n, const int *nh,
                                         const double *rin, const double *zin,
                                         const double *win, double *rout, double *zout,
                                         double *wout, double *pows, accel_rderivs *derivs);
ACCEL_HIDDEN void accel_set_error(const char *fmt, ...);
/* sig[i] = candidate_sigma(pow[i], numharm, numindep) as FP32, over host
 * threads when n is large (each value a pure function of its input). */
ACCEL_HIDDEN void accel_sigmas(const float *pow, long long n, int numharm, double numindep,
                               float *sig);

/* One candidate of the ACCEL text tables (src/accel_utils.c:666-979). */
typedef struct {
    float power, sigma;
    int numharm;
    double r, z, w;
    const accel_rderivs *derivs;       /* numharm entries */
    const double *hirs, *hizs, *hiws;  /* numharm entries */
} accel_out_cand;

/* The fundamentals table (src/accel_utils.c:666-842) and the harmonics
 * table (:844-972) of n candidates, each followed by two newlines. */
ACCEL_HIDDEN int accel_out_fundamentals(FILE *f, const accel_out_cand *c,
                                        const accel_fourierprops *props, int n, double T,
                                        int jerk);
ACCEL_HIDDEN int accel_out_harmonics(FILE *f, const accel_out_cand *c, int n, double T,
                                     int jerk);
/* Appends a file's bytes (the .inf) to f. */
ACCEL_HIDDEN void accel_out_append_file(FILE *f, const char *path);

#ifdef __cplusplus
}
#endif
#endif
