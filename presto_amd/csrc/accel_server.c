/*
 * accel_server.c -- a request loop over the accelsearch path on one GPU.
 *
 * Requests (a time series plus accelsearch parameters) are queued as they
 * arrive.  Worker threads admit them in arrival order whenever the device
 * memory they need (accel_search_footprint) fits under the budget next to
 * the requests already running.  While the oldest request waits for
 * memory, a later one may start beside it (backfill) if it fits now and
 * the requests started that way, together with the oldest, fit the budget:
 * the oldest then waits only for the requests that were running when it
 * reached the head of the queue, and small requests do not queue behind a
 * large one.  Each running request owns a search handle
 * and its HIP stream, so requests overlap on the device: one request's
 * upload, FFT and de-reddening, host replay and refinement run beside the
 * others' correlation and harmonic sums.  Idle handles stay cached (their
 * templates resident) and are reused by requests with equal parameters,
 * or destroyed, least recently used first, when memory is needed.
 *
 * No counterpart in the reference, which runs one accelsearch process per
 * input file (src/accelsearch.c:82-313); each request's results are those
 * of such a run.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "presto_accel.h"
#include "agpu.h"
#include "accel_internal.h"

#define SRV_MAXW 64
#define SRV_SCAN 256   /* queued requests looked at for a backfill */

typedef struct {
    accel_params p;
    float *series;
    long long n;
    int optimize;
    long long bytes;
    int state;                  /* 0 queued, 1 running, 2 done */
    int backfill;               /* started while an older request waited */
    int status;
    double t_submit, t_start, t_done;
    double ms_evict, ms_create, ms_work;   /* inside the service time */
    double ms_prep, ms_search, ms_refine;  /* inside ms_work */
    int ncands;
    accel_cand *cands;
    accel_fourierprops *props;  /* when optimized */
    char err[256];
} srv_req;

typedef struct {
    accel_params p;
    accel_search *s;
    long long bytes;
    int busy;
    long long id;
    double last_used;
} srv_handle;

struct accel_server {
    int device, nworkers;
    long long budget, in_use;   /* bytes of all live handles */
    pthread_mutex_t mu;
    pthread_cond_t cv;
    pthread_t th[SRV_MAXW];
    int shutdown;
    srv_req *req;
    int nreq, capreq, head;     /* queue = tickets head .. nreq-1 in state 0 */
    long long backfill_bytes;   /* footprints of running backfilled requests */
    srv_handle *hd;
    int nhd, caphd;
    long long next_id;
};

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* With the lock held: a handle for params p, reusing an idle cached one or
 * creating one when memory allows (evicting idle handles LRU).  Returns the
 * handle index, -1 if the request must wait, -2 if it can never fit. */
static int take_handle(accel_server *sv, const accel_params *p, long long bytes,
                       double *ms_evict)
{
    for (int i = 0; i < sv->nhd; i++)
        if (!sv->hd[i].busy && !memcmp(&sv->hd[i].p, p, sizeof(*p))) {
            sv->hd[i].busy = 1;
            return i;
        }
    if (bytes > sv->budget)
        return -2;
    while (sv->in_use + bytes > sv->budget) {
        int lru = -1;
        for (int i = 0; i < sv->nhd; i++)
            if (!sv->hd[i].busy && (lru < 0 || sv->hd[i].last_used < sv->hd[lru].last_used))
                lru = i;
        if (lru < 0)
            return -1;   /* running requests hold the memory: wait */
        double t0 = now_s();
        accel_search_destroy(sv->hd[lru].s);
        *ms_evict += 1e3 * (now_s() - t0);
        sv->in_use -= sv->hd[lru].bytes;
        sv->hd[lru] = sv->hd[--sv->nhd];
    }
    if (sv->nhd == sv->caphd) {
        sv->caphd = sv->caphd ? 2 * sv->caphd : 16;
        sv->hd = (srv_handle *) realloc(sv->hd, sizeof(srv_handle) * sv->caphd);
    }
    srv_handle *h = &sv->hd[sv->nhd];
    memset(h, 0, sizeof(*h));
    h->p = *p;
    h->bytes = bytes;
    h->busy = 1;
    h->id = ++sv->next_id;
    sv->in_use += bytes;   /* reserved now, created outside the lock */
    return sv->nhd++;
}

static void run_request(accel_server *sv, srv_req *r, accel_search *s)
{
    r->status = -1;
    double t0 = now_s();
    if (accel_search_set_series(s, r->series, r->n)) {
        snprintf(r->err, sizeof(r->err), "%s", accel_last_error());
        return;
    }
    double t1 = now_s();
    r->ms_prep = 1e3 * (t1 - t0);
    if (accel_search_run(s)) {
        snprintf(r->err, sizeof(r->err), "%s", accel_last_error());
        return;
    }
    double t2 = now_s();
    r->ms_search = 1e3 * (t2 - t1);
    int n = accel_search_ncands(s);
    r->ncands = n;
    r->cands = (accel_cand *) malloc(sizeof(accel_cand) * (n ? n : 1));
    accel_search_cands(s, r->cands, n);
    if (r->optimize && n) {
        r->props = (accel_fourierprops *) malloc(sizeof(accel_fourierprops) * n);
        if (accel_search_optimize(s) || accel_search_fund_props(s, r->props, n) != n) {
            snprintf(r->err, sizeof(r->err), "%s", accel_last_error());
            return;
        }
    }
    r->ms_refine = 1e3 * (now_s() - t2);
    r->status = 0;
    (void) sv;
}

static void *worker(void *arg)
{
    accel_server *sv = (accel_server *) arg;
    pthread_mutex_lock(&sv->mu);
    for (;;) {
        /* the oldest queued request, if its memory can be had */
        while (sv->head < sv->nreq && sv->req[sv->head].state != 0)
            sv->head++;
        if (sv->head >= sv->nreq) {
            if (sv->shutdown)
                break;
            pthread_cond_wait(&sv->cv, &sv->mu);
            continue;
        }
        int ti = sv->head;
        double ms_evict = 0.0;
        int hi = take_handle(sv, &sv->req[ti].p, sv->req[ti].bytes, &ms_evict);
        sv->req[ti].ms_evict += ms_evict;
        if (hi == -1) {
            /* the oldest request waits for memory: a later one that fits
             * now may start if it and the other backfilled requests leave
             * room for the oldest */
            const long long need = sv->req[sv->head].bytes;
            ti = -1;
            for (int k = sv->head + 1; k < sv->nreq && k <= sv->head + SRV_SCAN; k++) {
                srv_req *q = &sv->req[k];
                if (q->state != 0 || sv->backfill_bytes + q->bytes + need > sv->budget)
                    continue;
                double me = 0.0;
                hi = take_handle(sv, &q->p, q->bytes, &me);
                q->ms_evict += me;
                if (hi >= 0) {
                    ti = k;
                    break;
                }
            }
            if (ti < 0) {
                pthread_cond_wait(&sv->cv, &sv->mu);
                continue;
            }
            sv->req[ti].backfill = 1;
            sv->backfill_bytes += sv->req[ti].bytes;
        }
        srv_req *r = &sv->req[ti];
        r->state = 1;
        r->t_start = now_s();
        if (ti == sv->head)
            sv->head++;
        if (hi == -2) {
            snprintf(r->err, sizeof(r->err), "request needs %lld bytes, budget %lld", r->bytes,
                     sv->budget);
            r->status = -1;
            r->state = 2;
            r->t_done = now_s();
            pthread_cond_broadcast(&sv->cv);
            continue;
        }
        accel_search *s = sv->hd[hi].s;
        accel_params p = sv->hd[hi].p;
        const long long hid = sv->hd[hi].id;
        /* accel_server_submit may move the request array once the lock is
         * released: work on a copy taken under the lock */
        srv_req local = *r;
        pthread_mutex_unlock(&sv->mu);
        /* the current device is per host thread */
        agpu_init(sv->device);
        double t0 = now_s();
        if (!s)
            s = accel_search_create_streams(&p, sv->device, 1);
        double t1 = now_s();
        if (s)
            run_request(sv, &local, s);
        else
            snprintf(local.err, sizeof(local.err), "%s", accel_last_error());
        /* a handle whose run failed may hold a half-rebuilt state: it is
         * destroyed rather than cached */
        if (s && local.status != 0) {
            accel_search_destroy(s);
            s = NULL;
        }
        double t2 = now_s();
        pthread_mutex_lock(&sv->mu);
        /* the handle array may have moved too: find this handle again */
        for (int i = 0; i < sv->nhd; i++)
            if (sv->hd[i].id == hid) {
                if (s) {
                    sv->hd[i].s = s;
                    sv->hd[i].busy = 0;
                    sv->hd[i].last_used = now_s();
                } else {
                    sv->in_use -= sv->hd[i].bytes;
                    sv->hd[i] = sv->hd[--sv->nhd];
                }
                break;
            }
        r = &sv->req[ti];
        r->ms_create = 1e3 * (t1 - t0);
        r->ms_work = 1e3 * (t2 - t1);
        r->ms_prep = local.ms_prep;
        r->ms_search = local.ms_search;
        r->ms_refine = local.ms_refine;
        r->status = local.status;
        r->ncands = local.ncands;
        r->cands = local.cands;
        r->props = local.props;
        memcpy(r->err, local.err, sizeof(r->err));
        free(r->series);
        r->series = NULL;
        if (r->backfill)
            sv->backfill_bytes -= r->bytes;
        r->t_done = now_s();
        r->state = 2;
        pthread_cond_broadcast(&sv->cv);
    }
    pthread_mutex_unlock(&sv->mu);
    return NULL;
}

accel_server *accel_server_create(int device, long long hbm_budget, int workers)
{
    if (workers < 1 || workers > SRV_MAXW || hbm_budget <= 0)
        return NULL;
    /* every handle of the loop lives in one arena: no device memory is
     * returned to the driver between requests */
    if (agpu_init(device) ||
        agpu_arena_init((size_t) hbm_budget + (size_t) hbm_budget / 20 + ((size_t) 1 << 30)))
        return NULL;
    accel_server *sv = (accel_server *) calloc(1, sizeof(accel_server));
    sv->device = device;
    sv->budget = hbm_budget;
    sv->nworkers = workers;
    pthread_mutex_init(&sv->mu, NULL);
    pthread_cond_init(&sv->cv, NULL);
    for (int i = 0; i < workers; i++)
        pthread_create(&sv->th[i], NULL, worker, sv);
    return sv;
}

int accel_server_submit(accel_server *sv, const accel_params *p, const float *series,
                        long long n, int optimize)
{
    long long bytes = accel_search_footprint(p);
    if (bytes < 0 || n <= 0)
        return -1;
    float *copy = (float *) malloc(sizeof(float) * (size_t) n);
    memcpy(copy, series, sizeof(float) * (size_t) n);
    pthread_mutex_lock(&sv->mu);
    if (sv->nreq == sv->capreq) {
        sv->capreq = sv->capreq ? 2 * sv->capreq : 64;
        sv->req = (srv_req *) realloc(sv->req, sizeof(srv_req) * sv->capreq);
    }
    const int t = sv->nreq++;
    srv_req *r = &sv->req[t];
    memset(r, 0, sizeof(*r));
    r->p = *p;
    r->series = copy;
    r->n = n;
    r->optimize = optimize;
    r->bytes = bytes;
    r->t_submit = now_s();
    pthread_cond_broadcast(&sv->cv);
    pthread_mutex_unlock(&sv->mu);
    return t;
}

int accel_server_wait(accel_server *sv, int ticket, double *latency_ms, double *service_ms)
{
    pthread_mutex_lock(&sv->mu);
    if (ticket < 0 || ticket >= sv->nreq) {
        pthread_mutex_unlock(&sv->mu);
        return -1;
    }
    while (sv->req[ticket].state != 2)
        pthread_cond_wait(&sv->cv, &sv->mu);
    srv_req *r = &sv->req[ticket];
    if (latency_ms)
        *latency_ms = 1e3 * (r->t_done - r->t_submit);
    if (service_ms)
        *service_ms = 1e3 * (r->t_done - r->t_start);
    int st = r->status;
    pthread_mutex_unlock(&sv->mu);
    return st;
}

int accel_server_timing(accel_server *sv, int ticket, double *ms9)
{
    pthread_mutex_lock(&sv->mu);
    int rc = -1;
    if (ticket >= 0 && ticket < sv->nreq && sv->req[ticket].state == 2) {
        srv_req *r = &sv->req[ticket];
        ms9[0] = 1e3 * (r->t_start - r->t_submit);   /* queued */
        ms9[1] = r->ms_evict;                         /* evicting cached handles */
        ms9[2] = r->ms_create;                        /* creating the handle */
        ms9[3] = r->ms_prep;                          /* upload, real FFT, de-reddening */
        ms9[4] = r->ms_search;                        /* search through the candidate list */
        ms9[5] = r->ms_refine;                        /* refinement and properties */
        ms9[6] = r->ms_work;                          /* the three above */
        ms9[7] = 1e3 * (r->t_done - r->t_start);      /* service */
        ms9[8] = (double) r->bytes;
        rc = 0;
    }
    pthread_mutex_unlock(&sv->mu);
    return rc;
}

int accel_server_cands(accel_server *sv, int ticket, accel_cand *out, int n)
{
    pthread_mutex_lock(&sv->mu);
    int rc = -1;
    if (ticket >= 0 && ticket < sv->nreq && sv->req[ticket].state == 2) {
        srv_req *r = &sv->req[ticket];
        rc = r->ncands;
        if (out && r->cands)
            memcpy(out, r->cands, sizeof(accel_cand) * (size_t) (n < rc ? n : rc));
    }
    pthread_mutex_unlock(&sv->mu);
    return rc;
}

int accel_server_props(accel_server *sv, int ticket, accel_fourierprops *out, int n)
{
    pthread_mutex_lock(&sv->mu);
    int rc = -1;
    if (ticket >= 0 && ticket < sv->nreq && sv->req[ticket].state == 2 &&
        sv->req[ticket].props) {
        srv_req *r = &sv->req[ticket];
        rc = r->ncands;
        if (out)
            memcpy(out, r->props, sizeof(accel_fourierprops) * (size_t) (n < rc ? n : rc));
    }
    pthread_mutex_unlock(&sv->mu);
    return rc;
}

long long accel_server_memory(accel_server *sv)
{
    pthread_mutex_lock(&sv->mu);
    long long b = sv->in_use;
    pthread_mutex_unlock(&sv->mu);
    return b;
}

void accel_server_destroy(accel_server *sv)
{
    if (!sv)
        return;
    pthread_mutex_lock(&sv->mu);
    sv->shutdown = 1;
    pthread_cond_broadcast(&sv->cv);
    pthread_mutex_unlock(&sv->mu);
    for (int i = 0; i < sv->nworkers; i++)
        pthread_join(sv->th[i], NULL);
    for (int i = 0; i < sv->nhd; i++)
        accel_search_destroy(sv->hd[i].s);
    for (int i = 0; i < sv->nreq; i++) {
        free(sv->req[i].series);
        free(sv->req[i].cands);
        free(sv->req[i].props);
    }
    free(sv->req);
    free(sv->hd);
    accel_template_cache_clear();
    agpu_arena_release();
    pthread_mutex_destroy(&sv->mu);
    pthread_cond_destroy(&sv->cv);
    free(sv);
}
