/*
 * Host-only driver for the request loop (accel_server.c) under
 * ThreadSanitizer or AddressSanitizer, built by tests/test_sanitize.py.
 * The search handle and the device layer are replaced by test doubles
 * defined here: a "search" sleeps for a time that grows with its
 * footprint and returns candidates computed from its series and
 * parameters, and creating / destroying handles books their footprint in
 * a live-memory counter.  The driver checks what the loop promises:
 *
 *   - every request finishes, each with the results of its own series and
 *     parameters, whatever else runs beside it;
 *   - the live handles never hold more than the budget;
 *   - a request larger than the budget fails and leaves nothing behind;
 *   - the request array grows while workers run (more than 64 requests);
 *   - while the oldest request waits for memory, a small later one starts
 *     beside it (backfill), and the oldest is not kept waiting by it.
 *
 * Exits 0 when every check holds; the sanitizers abort on the first error.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>
#include "presto_accel.h"
#include "agpu.h"

#define CHECK(c)                                                           \
    do {                                                                   \
        if (!(c)) {                                                        \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            exit(1);                                                       \
        }                                                                  \
    } while (0)

/* ---- test doubles of the search handle and the device layer ---- */

struct accel_search {
    accel_params p;
    long long bytes;
    double sum;
    int ncands;
};

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static long long g_live, g_live_max, g_created, g_destroyed;
static int g_us_per_mib = 20;   /* "search" time per MiB of footprint */

long long accel_search_footprint(const accel_params *p)
{
    /* 1 MiB per 1024 samples, times 2 with -wmax */
    return ((long long) p->N << 10) * (p->wmaxP ? 2 : 1);
}

accel_search *accel_search_create(const accel_params *p, int device)
{
    (void) device;
    accel_search *s = (accel_search *) calloc(1, sizeof(*s));
    s->p = *p;
    s->bytes = accel_search_footprint(p);
    pthread_mutex_lock(&g_mu);
    g_live += s->bytes;
    if (g_live > g_live_max)
        g_live_max = g_live;
    g_created++;
    pthread_mutex_unlock(&g_mu);
    return s;
}

accel_search *accel_search_create_streams(const accel_params *p, int device, int nstreams)
{
    (void) nstreams;
    return accel_search_create(p, device);
}

void accel_search_destroy(accel_search *s)
{
    pthread_mutex_lock(&g_mu);
    g_live -= s->bytes;
    g_destroyed++;
    pthread_mutex_unlock(&g_mu);
    free(s);
}

int accel_search_set_series(accel_search *s, const float *series, long long n)
{
    double sum = 0.0;
    for (long long i = 0; i < n; i++)
        sum += series[i];
    s->sum = sum;
    return 0;
}

int accel_search_run(accel_search *s)
{
    usleep((useconds_t) (g_us_per_mib * (s->bytes >> 20)));
    s->ncands = 1 + (int) ((long long) s->p.N % 3);
    return 0;
}

int accel_search_ncands(const accel_search *s) { return s->ncands; }

static accel_cand expected_cand(const accel_params *p, double sum, int i)
{
    accel_cand c;
    memset(&c, 0, sizeof(c));
    c.power = (float) (sum + i);
    c.sigma = (float) p->zmax;
    c.numharm = p->numharm;
    c.r = p->N + i;
    c.z = p->zmax;
    c.w = p->wmaxP ? p->wmax : 0;
    return c;
}

int accel_search_cands(const accel_search *s, accel_cand *out, int n)
{
    for (int i = 0; i < n && i < s->ncands; i++)
        out[i] = expected_cand(&s->p, s->sum, i);
    return n < s->ncands ? n : s->ncands;
}

int accel_search_optimize(accel_search *s)
{
    usleep(50);
    (void) s;
    return 0;
}

int accel_search_fund_props(const accel_search *s, accel_fourierprops *props, int n)
{
    for (int i = 0; i < n && i < s->ncands; i++) {
        memset(&props[i], 0, sizeof(props[i]));
        props[i].r = s->p.N + i;
        props[i].pow = (float) s->sum;
    }
    return n < s->ncands ? n : s->ncands;
}

const char *accel_last_error(void) { return "test double"; }
void accel_template_cache_clear(void) {}
int agpu_init(int device) { (void) device; return 0; }
int agpu_arena_init(size_t bytes) { (void) bytes; return 0; }
int agpu_arena_release(void) { return 0; }

/* ---- the checks ---- */

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return 1e3 * ts.tv_sec + 1e-6 * ts.tv_nsec;
}

static accel_params params(double N, int zmax, int wmax, int numharm)
{
    accel_params p;
    memset(&p, 0, sizeof(p));
    p.N = N;
    p.dt = 1e-4;
    p.zmax = zmax;
    p.wmax = wmax;
    p.wmaxP = wmax > 0;
    p.numharm = numharm;
    p.sigma = 2.0f;
    return p;
}

static float *series(long long n, unsigned seed, double *sum)
{
    float *x = (float *) malloc(sizeof(float) * (size_t) n);
    double s = 0.0;
    for (long long i = 0; i < n; i++) {
        seed = seed * 1103515245u + 12345u;
        x[i] = (float) ((seed >> 16) & 1023) / 1024.0f;
        s += x[i];
    }
    *sum = s;
    return x;
}

static void check_results(accel_server *sv, int t, const accel_params *p, double sum,
                          int optimized)
{
    double lat = 0.0, svc = 0.0;
    CHECK(accel_server_wait(sv, t, &lat, &svc) == 0);
    CHECK(lat >= svc && svc >= 0.0);
    accel_cand c[8];
    const int n = accel_server_cands(sv, t, c, 8);
    CHECK(n == 1 + (int) ((long long) p->N % 3));
    for (int i = 0; i < n; i++) {
        const accel_cand e = expected_cand(p, sum, i);
        CHECK(memcmp(&c[i], &e, sizeof(e)) == 0);
    }
    accel_fourierprops props[8];
    if (optimized) {
        CHECK(accel_server_props(sv, t, props, 8) == n);
        for (int i = 0; i < n; i++)
            CHECK(props[i].r == p->N + i && props[i].pow == (float) sum);
    } else {
        CHECK(accel_server_props(sv, t, props, 8) == -1);   /* not refined */
    }
    double ms9[9];
    CHECK(accel_server_timing(sv, t, ms9) == 0);
    CHECK(ms9[8] == (double) accel_search_footprint(p));
}

/* 200 requests of mixed sizes and parameters through 4 workers, submitted
 * while the workers run */
static void mixed(void)
{
    /* the largest request (2^18 samples with -wmax) takes most of it, so
     * requests wait for memory, handles are evicted and later ones backfill */
    const long long budget = 600LL << 20;
    accel_server *sv = accel_server_create(0, budget, 4);
    CHECK(sv != NULL);
    enum { NR = 200 };
    static accel_params ps[NR];
    static double sums[NR];
    static int tickets[NR];
    for (int i = 0; i < NR; i++) {
        const double N = (double) ((1 << 12) << (i % 7)) + 2 * i;
        ps[i] = params(N, (i % 3) * 50, (i % 5 == 0) ? 100 : 0, 8 << (i % 2));
        float *x = series((long long) N, 77u + (unsigned) i, &sums[i]);
        tickets[i] = accel_server_submit(sv, &ps[i], x, (long long) N, i % 4 != 0);
        free(x);
        CHECK(tickets[i] == i);
    }
    for (int i = 0; i < NR; i++)
        check_results(sv, tickets[i], &ps[i], sums[i], i % 4 != 0);
    CHECK(accel_server_memory(sv) <= budget);
    accel_server_destroy(sv);
    CHECK(g_live == 0);
    CHECK(g_live_max <= budget);
}

/* a request above the budget fails cleanly */
static void too_large(void)
{
    accel_params p = params(1 << 16, 200, 0, 8);
    const long long budget = accel_search_footprint(&p) / 2;
    accel_server *sv = accel_server_create(0, budget, 2);
    double sum;
    float *x = series(1 << 16, 5u, &sum);
    const int t = accel_server_submit(sv, &p, x, 1 << 16, 1);
    free(x);
    double lat, svc;
    CHECK(accel_server_wait(sv, t, &lat, &svc) == -1);
    CHECK(accel_server_memory(sv) == 0);
    accel_server_destroy(sv);
    CHECK(g_live == 0);
}

/* two large requests that cannot share the budget, then a small one: the
 * small one starts while the second large one waits */
static void backfill(void)
{
    /* 1 GiB each: 300 ms of "search", long beside the submissions even
     * under a sanitizer */
    g_us_per_mib = 300;
    accel_params big = params(1 << 20, 200, 0, 16), big2 = params((1 << 20) + 2, 200, 0, 16);
    accel_params small = params(1 << 12, 0, 0, 2);
    const long long fb = accel_search_footprint(&big), fs = accel_search_footprint(&small);
    const long long budget = fb + 2 * fs;
    accel_server *sv = accel_server_create(0, budget, 4);
    double sa, sh, ss;
    float *xa = series(1 << 20, 11u, &sa), *xh = series((1 << 20) + 2, 12u, &sh);
    float *xs = series(1 << 12, 13u, &ss);
    const double t0 = now_ms();
    const int ta = accel_server_submit(sv, &big, xa, 1 << 20, 1);
    usleep(2000);   /* the first large request is running */
    const int th = accel_server_submit(sv, &big2, xh, (1 << 20) + 2, 1);
    const double t1 = now_ms();
    const int ts = accel_server_submit(sv, &small, xs, 1 << 12, 1);
    free(xa);
    free(xh);
    free(xs);
    check_results(sv, ta, &big, sa, 1);
    check_results(sv, th, &big2, sh, 1);
    check_results(sv, ts, &small, ss, 1);
    double ma[9], mh[9], ms[9], lat_a, svc_a;
    CHECK(accel_server_timing(sv, ta, ma) == 0 && accel_server_timing(sv, th, mh) == 0 &&
          accel_server_timing(sv, ts, ms) == 0);
    CHECK(accel_server_wait(sv, ta, &lat_a, &svc_a) == 0);
    /* the small request started while the first large one still ran (in
     * arrival order it would have waited for it and for the second) */
    if (getenv("SRV_DEBUG"))
        fprintf(stderr, "t1-t0 %.2f ms, small queued %.2f, large 1 latency %.2f (queued %.2f service %.2f), large 2 queued %.2f\n", t1 - t0, ms[0], lat_a, ma[0], ma[7], mh[0]);
    CHECK(t1 + ms[0] < t0 + lat_a);
    CHECK(ms[0] < mh[0]);
    accel_server_destroy(sv);
    g_us_per_mib = 20;
    CHECK(g_live == 0);
    CHECK(g_live_max <= budget);
}

int main(void)
{
    mixed();
    too_large();
    g_live_max = 0;
    backfill();
    printf("request loop checks passed (%lld handles created, %lld destroyed)\n", g_created,
           g_destroyed);
    return 0;
}
