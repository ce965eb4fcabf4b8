/*
 * accel_batch.c -- many DM-trial spectra through the search on one GPU.
 *
 * The reference runs accelsearch once per spectrum (src/accelsearch.c:
 * 82-313).  Here a batch keeps a pool of search handles keyed by search
 * parameters (templates, plane, windows resident in HBM and reused by
 * every spectrum of the same geometry) and pipelines the spectra: while
 * spectrum i+1 is uploaded and searched on one handle's stream, the hits
 * of spectrum i are replayed through the candidate logic, its candidates
 * are refined on the GPU on the other handle's stream (a few workgroups,
 * concurrent with the search) and its files are written.  Every spectrum
 * gets exactly the candidates, refinement and files of a separate run.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "presto_accel.h"

#define POOL_MAX 64

typedef struct {
    accel_params p;
    accel_search *s;
    int last_item;   /* item whose results the handle still holds, or -1 */
} pooled;

struct accel_batch {
    int device;
    int n;
    pooled pool[POOL_MAX];
    char err[512];
};

accel_batch *accel_batch_create(int device)
{
    accel_batch *b = (accel_batch *) calloc(1, sizeof(accel_batch));
    b->device = device;
    return b;
}

void accel_batch_destroy(accel_batch *b)
{
    if (!b)
        return;
    for (int i = 0; i < b->n; i++)
        accel_search_destroy(b->pool[i].s);
    free(b);
}

int accel_batch_handles(const accel_batch *b)
{
    return b->n;
}

/* A handle for params p that is not holding the results of item `busy`. */
static pooled *get_handle(accel_batch *b, const accel_params *p, int busy)
{
    for (int i = 0; i < b->n; i++)
        if (!memcmp(&b->pool[i].p, p, sizeof(*p)) && b->pool[i].last_item != busy)
            return &b->pool[i];
    if (b->n == POOL_MAX)
        return NULL;
    pooled *q = &b->pool[b->n];
    q->s = accel_search_create(p, b->device);
    if (!q->s)
        return NULL;
    q->p = *p;
    q->last_item = -1;
    b->n++;
    return q;
}

/* Replay, refinement, outputs and result copy of one searched item. */
static int finish(pooled *q, accel_batch_item *it)
{
    if (accel_search_collect(q->s))
        return -1;
    int n = accel_search_ncands(q->s);
    it->ncands = n;
    if (it->optimize && n) {
        if (accel_search_optimize(q->s))
            return -1;
        if (it->accelnm &&
            accel_search_write_outputs(q->s, it->accelnm, it->candnm, it->inf_path))
            return -1;
    }
    if (it->cands && it->maxcands > 0)
        accel_search_cands(q->s, it->cands, n < it->maxcands ? n : it->maxcands);
    float ms[5];
    accel_search_timings(q->s, ms);
    it->gpu_ms = ms[4];
    return 0;
}

int accel_batch_run(accel_batch *b, accel_batch_item *items, int n)
{
    pooled *prev = NULL;
    int previ = -1, rc = 0;
    for (int i = 0; i < n; i++) {
        accel_batch_item *it = &items[i];
        it->status = 0;
        if (it->params.numbins <= 0 || (!it->fft && !it->d_series)) {
            it->status = -1;
            rc = -1;
            continue;
        }
        pooled *q = get_handle(b, &it->params, previ);
        if (!q ||
            (it->fft ? accel_search_set_spectrum(q->s, it->fft)
                     : accel_search_set_device_series(q->s, it->d_series, it->nseries)) ||
            accel_search_launch(q->s)) {
            it->status = -1;
            rc = -1;
            continue;
        }
        q->last_item = i;
        /* the previous spectrum's host work and refinement overlap this search */
        if (prev && finish(prev, &items[previ])) {
            items[previ].status = -1;
            rc = -1;
        }
        prev = q;
        previ = i;
    }
    if (prev && finish(prev, &items[previ])) {
        items[previ].status = -1;
        rc = -1;
    }
    for (int i = 0; i < b->n; i++)
        b->pool[i].last_item = -1;
    return rc;
}
