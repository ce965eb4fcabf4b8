/*
 * accel_sift.c -- sifting of accelsearch candidates across DM trials.
 *
 * The behaviour is that of PRESTO's sifting module
 * (python/presto/sifting.py:14-1206) and its example driver
 * (examplescripts/ACCEL_sift.py): ACCEL text files are read back, each
 * candidate is re-scored from its refined harmonic powers, obviously bad
 * candidates are set aside by simple rules, detections of one signal in
 * several DM trials are merged into the strongest, candidates whose DM
 * behaviour is not a pulsar's are set aside, and so are harmonics of
 * stronger candidates.
 *
 * Here it is a C candidate set: every candidate lives in exactly one list
 * (the good list, the duplicates, or one of eight reject lists), and each
 * operation moves candidates between lists in a documented order, so that
 * every output (the candidate list text, the report, the summary) follows
 * from the inputs alone.
 */
#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "presto_accel.h"
#include "accel_internal.h"

typedef struct {
    double dm, snr, sigma;
} hit_t;

typedef struct {
    char path[512];       /* as given */
    char file[256];       /* its last path component */
    char dmstr[32];
    int candnum, numharm;
    double sigma, ipow, cpow, r, z, T, dm, snr;
    double f, p;          /* r / T (Hz) and 1 / f (s), rounded as the sifter's fields are */
    double *pows, *phs;   /* numharm each (phs: radians) */
    hit_t *hits;
    int nhits, caphits;
    char note[512];
    long pos;             /* sort tie-break: position before a sort */
} cand_t;

typedef struct {
    cand_t **v;
    int n, cap;
} clist;

struct accel_sift {
    clist list[ACCEL_SIFT_NLISTS];   /* ACCEL_SIFT_GOOD, _DUPLICATE, _BAD_... */
    int track_bad, track_dupes;
};

/* ---- lists ---- */

static int cl_push(clist *l, cand_t *c)
{
    if (l->n == l->cap) {
        int cap = l->cap ? 2 * l->cap : 64;
        cand_t **v = (cand_t **) realloc(l->v, sizeof(cand_t *) * (size_t) cap);
        if (!v)
            return -1;
        l->v = v;
        l->cap = cap;
    }
    l->v[l->n++] = c;
    return 0;
}

/* removes entry i, keeping the order of the rest */
static cand_t *cl_take(clist *l, int i)
{
    cand_t *c = l->v[i];
    memmove(l->v + i, l->v + i + 1, sizeof(cand_t *) * (size_t) (l->n - i - 1));
    l->n--;
    return c;
}

static void cand_free(cand_t *c)
{
    if (!c)
        return;
    free(c->pows);
    free(c->phs);
    free(c->hits);
    free(c);
}

static int add_hits(cand_t *c, const hit_t *h, int n)
{
    if (c->nhits + n > c->caphits) {
        int cap = c->caphits ? c->caphits : 4;
        while (cap < c->nhits + n)
            cap *= 2;
        hit_t *v = (hit_t *) realloc(c->hits, sizeof(hit_t) * (size_t) cap);
        if (!v)
            return -1;
        c->hits = v;
        c->caphits = cap;
    }
    memcpy(c->hits + c->nhits, h, sizeof(hit_t) * (size_t) n);
    c->nhits += n;
    return 0;
}

/* Moves good candidate i to list `to`; a removed candidate that the set
 * does not keep (untracked) is freed. */
static void move_good(accel_sift *s, int i, int to)
{
    cand_t *c = cl_take(&s->list[ACCEL_SIFT_GOOD], i);
    const int keep = to == ACCEL_SIFT_DUPLICATE ? s->track_dupes : s->track_bad;
    if (!keep || cl_push(&s->list[to], c))
        cand_free(c);
}

static void set_note(cand_t *c, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->note, sizeof(c->note), fmt, ap);
    va_end(ap);
}

/* ---- stable sorts (ties keep their current order) ---- */

static int by_r(const void *a, const void *b)
{
    const cand_t *x = *(cand_t *const *) a, *y = *(cand_t *const *) b;
    if (x->r != y->r)
        return x->r < y->r ? -1 : 1;
    return (x->pos > y->pos) - (x->pos < y->pos);
}

static int by_sigma_down(const void *a, const void *b)
{
    const cand_t *x = *(cand_t *const *) a, *y = *(cand_t *const *) b;
    if (x->sigma != y->sigma)
        return x->sigma > y->sigma ? -1 : 1;
    return (x->pos > y->pos) - (x->pos < y->pos);
}

static void stable_sort(clist *l, int (*cmp)(const void *, const void *))
{
    for (int i = 0; i < l->n; i++)
        l->v[i]->pos = i;
    qsort(l->v, (size_t) l->n, sizeof(cand_t *), cmp);
}

/* sorts a candidate's hits by DM, equal DMs keeping their order */
static void sort_hits_dm(cand_t *c)
{
    /* insertion sort: hit lists are short, and it is stable */
    for (int i = 1; i < c->nhits; i++) {
        hit_t h = c->hits[i];
        int j = i - 1;
        while (j >= 0 && c->hits[j].dm > h.dm) {
            c->hits[j + 1] = c->hits[j];
            j--;
        }
        c->hits[j + 1] = h;
    }
}

/* ---- scoring ---- */

/* S/N of the harmonics: sum of sqrt(power - 1) over harmonics above the
 * mean noise power of 1 (python/presto/sifting.py:166-171). */
static double harm_snr(const double *pows, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; i++)
        s += pows[i] > 1.0 ? sqrt(pows[i] - 1.0) : 0.0;
    return s;
}

/* The refined score of a candidate: incoherent power = sum of the refined
 * harmonic powers, sigma for one trial, S/N, and the candidate as its own
 * single hit (python/presto/sifting.py:1137-1170). */
static int rescore(cand_t *c)
{
    double ip = 0.0;
    for (int i = 0; i < c->numharm; i++)
        ip += c->pows[i];
    c->ipow = ip;
    c->sigma = accel_candidate_sigma(ip, c->numharm, 1.0);
    c->snr = harm_snr(c->pows, c->numharm);
    c->nhits = 0;
    hit_t h = {c->dm, c->snr, c->sigma};
    return add_hits(c, &h, 1);
}

/* ---- construction ---- */

accel_sift *accel_sift_create(int track_bad, int track_dupes)
{
    accel_sift *s = (accel_sift *) calloc(1, sizeof(accel_sift));
    if (!s) {
        accel_set_error("accel_sift_create: out of host memory");
        return NULL;
    }
    s->track_bad = track_bad != 0;
    s->track_dupes = track_dupes != 0;
    return s;
}

void accel_sift_destroy(accel_sift *s)
{
    if (!s)
        return;
    for (int l = 0; l < ACCEL_SIFT_NLISTS; l++) {
        for (int i = 0; i < s->list[l].n; i++)
            cand_free(s->list[l].v[i]);
        free(s->list[l].v);
    }
    free(s);
}

static void split_path(cand_t *c, const char *path)
{
    snprintf(c->path, sizeof(c->path), "%s", path);
    const char *slash = strrchr(path, '/');
    snprintf(c->file, sizeof(c->file), "%s", slash ? slash + 1 : path);
}

/* The DM of an ACCEL file: the first "DM<digits>.<2 digits>" in its path
 * (python/presto/sifting.py:42, 1118). */
static int dm_from_path(const char *path, char *dmstr, size_t len, double *dm)
{
    for (const char *p = strstr(path, "DM"); p; p = strstr(p + 1, "DM")) {
        const char *q = p + 2;
        if (!isdigit((unsigned char) *q))
            continue;
        while (isdigit((unsigned char) *q))
            q++;
        if (q[0] != '.' || !isdigit((unsigned char) q[1]) || !isdigit((unsigned char) q[2]))
            continue;
        size_t n = (size_t) (q + 3 - (p + 2));
        if (n + 1 > len)
            return -1;
        memcpy(dmstr, p + 2, n);
        dmstr[n] = '\0';
        *dm = strtod(dmstr, NULL);
        return 0;
    }
    return -1;
}

int accel_sift_add(accel_sift *s, const accel_sift_cand *in, const double *harm_pows,
                   const double *harm_phases, const double *hits, int nhits)
{
    if (!s || !in || in->numharm < 1 || in->numharm > ACCEL_MAXHARM || nhits < 0 ||
        !(in->T > 0.0)) {
        accel_set_error("accel_sift_add: bad candidate");
        return -1;
    }
    cand_t *c = (cand_t *) calloc(1, sizeof(cand_t));
    if (!c)
        goto oom;
    split_path(c, in->path);
    snprintf(c->dmstr, sizeof(c->dmstr), "%s", in->dmstr);
    c->dm = strtod(in->dmstr, NULL);
    c->candnum = in->candnum;
    c->numharm = in->numharm;
    c->sigma = in->sigma;
    c->ipow = in->ipow;
    c->cpow = in->cpow;
    c->r = in->r;
    c->z = in->z;
    c->T = in->T;
    c->f = c->r / c->T;
    c->p = 1.0 / c->f;
    c->snr = in->snr;
    c->pows = (double *) calloc((size_t) c->numharm, sizeof(double));
    c->phs = (double *) calloc((size_t) c->numharm, sizeof(double));
    if (!c->pows || !c->phs)
        goto oom;
    for (int i = 0; i < c->numharm; i++) {
        c->pows[i] = harm_pows ? harm_pows[i] : 0.0;
        c->phs[i] = harm_phases ? harm_phases[i] : 0.0;
    }
    for (int i = 0; i < nhits; i++) {
        hit_t h = {hits[3 * i], hits[3 * i + 1], hits[3 * i + 2]};
        if (add_hits(c, &h, 1))
            goto oom;
    }
    if (cl_push(&s->list[ACCEL_SIFT_GOOD], c))
        goto oom;
    return 0;
oom:
    cand_free(c);
    accel_set_error("accel_sift_add: out of host memory");
    return -1;
}

/* "value(err)" or "value(err)x10^e" (an exponent cut off by the column
 * width counts as 5): python/presto/sifting.py:124-132. */
double accel_sift_parse_power(const char *txt)
{
    double v = strtod(txt, NULL);
    const char *e = strchr(txt, '^');
    if (e) {
        char *end;
        double ex = strtod(e + 1, &end);
        if (end == e + 1)
            ex = 5.0;
        v *= pow(10.0, ex);
    }
    return v;
}

#define MAXTOK 24
static int tokens(char *line, char **tok)
{
    int n = 0;
    for (char *p = strtok(line, " \t\r\n"); p && n < MAXTOK; p = strtok(NULL, " \t\r\n"))
        tok[n++] = p;
    return n;
}

/* Reads the candidates of one ACCEL file (src/accel_utils.c:629-979 wrote
 * it): the fundamentals table gives candnum, sigma, powers, numharm, r and
 * z; the k-th harmonics table the k-th candidate's refined harmonic powers
 * and phases; the appended .inf the observation length.  Each candidate is
 * then re-scored from its harmonics.  Returns the candidates read or -1. */
int accel_sift_read(accel_sift *s, const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f) {
        accel_set_error("accel_sift_read: cannot open %s", path);
        return -1;
    }
    char dmstr[32];
    double dm = 0.0;
    if (dm_from_path(path, dmstr, sizeof(dmstr), &dm)) {
        fclose(f);
        accel_set_error("accel_sift_read: no DM<value> in the name %s", path);
        return -1;
    }
    /* the observation length from the appended .inf */
    char line[4096];
    long long numsamp = -1;
    double dt = -1.0;
    while (fgets(line, sizeof(line), f)) {
        if (!strncmp(line, " Number of bins in the time series", 34)) {
            char *last = NULL, *tok[MAXTOK];
            int n = tokens(line, tok);
            if (n)
                last = tok[n - 1];
            numsamp = last ? atoll(last) : -1;
        } else if (!strncmp(line, " Width of each time series bin (sec)", 36)) {
            char *tok[MAXTOK];
            int n = tokens(line, tok);
            dt = n ? strtod(tok[n - 1], NULL) : -1.0;
        }
    }
    if (numsamp <= 0 || !(dt > 0.0)) {
        fclose(f);
        accel_set_error("accel_sift_read: %s carries no .inf length and sample time", path);
        return -1;
    }
    const double T = (double) numsamp * dt;
    rewind(f);
    clist *good = &s->list[ACCEL_SIFT_GOOD];
    const int first = good->n;
    int ntab = 0;           /* harmonic tables started */
    cand_t *cur = NULL;     /* candidate whose harmonic rows are being read */
    int cur_h = 0;
    int rc = 0;
    while (fgets(line, sizeof(line), f)) {
        char *tok[MAXTOK];
        if (isdigit((unsigned char) line[0])) {
            int n = tokens(line, tok);
            if (n < 10) {
                rc = -1;
                break;
            }
            accel_sift_cand in;
            memset(&in, 0, sizeof(in));
            snprintf(in.path, sizeof(in.path), "%s", path);
            snprintf(in.dmstr, sizeof(in.dmstr), "%s", dmstr);
            in.candnum = atoi(tok[0]);
            in.sigma = strtod(tok[1], NULL);
            in.ipow = strtod(tok[2], NULL);
            in.cpow = strtod(tok[3], NULL);
            in.numharm = atoi(tok[4]);
            in.r = strtod(tok[7], NULL);
            in.z = strtod(tok[9], NULL);
            in.T = T;
            if (accel_sift_add(s, &in, NULL, NULL, NULL, 0)) {
                rc = -1;
                break;
            }
            /* candidate numbers run 1, 2, ... in file order; beyond 9999
             * the column is too narrow, so the position is the number */
            cand_t *c = good->v[good->n - 1];
            c->candnum = good->n - first;
            cur = NULL;
        } else if (line[0] == ' ' && isdigit((unsigned char) line[1])) {
            int n = tokens(line, tok);
            if (first + ntab >= good->n || n < 10) {
                rc = -1;
                break;
            }
            cur = good->v[first + ntab++];
            cur->pows[0] = accel_sift_parse_power(tok[3]);
            cur->phs[0] = strtod(tok[9], NULL);
            cur_h = 1;
            if (cur_h == cur->numharm) {
                if (rescore(cur)) {
                    rc = -1;
                    break;
                }
                cur = NULL;
            }
        } else if (cur) {
            int n = tokens(line, tok);
            if (n < 9) {
                rc = -1;
                break;
            }
            cur->pows[cur_h] = accel_sift_parse_power(tok[2]);
            cur->phs[cur_h] = strtod(tok[8], NULL);
            if (++cur_h == cur->numharm) {
                if (rescore(cur)) {
                    rc = -1;
                    break;
                }
                cur = NULL;
            }
        }
    }
    fclose(f);
    if (rc == 0 && (ntab != good->n - first || cur)) {
        accel_set_error("accel_sift_read: %s: %d candidates but %d harmonic tables", path,
                        good->n - first, ntab);
        rc = -1;
    } else if (rc) {
        accel_set_error("accel_sift_read: %s: a table line is malformed", path);
    }
    if (rc) {
        while (good->n > first)
            cand_free(cl_take(good, good->n - 1));
        return -1;
    }
    return good->n - first;
}

/* ---- rejection (python/presto/sifting.py:504-658) ---- */

/* The first rule of `rules` (ACCEL_SIFT_R_* bits, tested in the order of
 * the bits) a candidate fails, writing its note; 0 if none. */
static int failing_rule(const cand_t *c, cand_t *cw, const accel_sift_params *p, unsigned rules)
{
    const double f = c->f, per = c->p;
    if ((rules & ACCEL_SIFT_R_LONGPERIOD) && per > p->long_period) {
        set_note(cw, "Period is too long (%g ms > %g ms)", per * 1000.0, p->long_period * 1000.0);
        return ACCEL_SIFT_BAD_LONGPERIOD;
    }
    if ((rules & ACCEL_SIFT_R_SHORTPERIOD) && per < p->short_period) {
        set_note(cw, "Period is too short (%g ms < %g ms)", per * 1000.0,
                 p->short_period * 1000.0);
        return ACCEL_SIFT_BAD_SHORTPERIOD;
    }
    if (rules & ACCEL_SIFT_R_KNOWNBIRDS) {
        for (int b = 0; b < p->nbirds_f; b++)
            if (fabs(f - p->birds_f[2 * b]) < p->birds_f[2 * b + 1]) {
                set_note(cw, "Freq (%.2f Hz) is within %g Hz of a known birdie centred at %.2f Hz",
                         f, p->birds_f[2 * b + 1], p->birds_f[2 * b]);
                return ACCEL_SIFT_BAD_KNOWNBIRDS;
            }
        for (int b = 0; b < p->nbirds_p; b++)
            if (fabs(per * 1000.0 - p->birds_p[2 * b]) < p->birds_p[2 * b + 1]) {
                set_note(cw, "Period (%.2f ms) is within %g ms of a known birdie centred at %.2f ms",
                         per * 1000.0, p->birds_p[2 * b + 1], p->birds_p[2 * b]);
                return ACCEL_SIFT_BAD_KNOWNBIRDS;
            }
    }
    if (rules & ACCEL_SIFT_R_THRESHOLD) {
        if (c->numharm == 1) {
            if (c->sigma < p->sigma_threshold && c->cpow < p->c_pow_threshold) {
                set_note(cw, "Only 1 harmonic and both sigma (%g < %g) and coherent power "
                             "(%g < %g) are too low.",
                         c->sigma, p->sigma_threshold, c->cpow, p->c_pow_threshold);
                return ACCEL_SIFT_BAD_THRESHOLD;
            }
        } else if (c->sigma < p->sigma_threshold) {
            set_note(cw, "%d harmonics and sigma (%g < %g) is too low.", c->numharm, c->sigma,
                     p->sigma_threshold);
            return ACCEL_SIFT_BAD_THRESHOLD;
        }
    }
    /* the strongest harmonic (the first of equals) and the next power */
    int imax = 0;
    for (int h = 1; h < c->numharm; h++)
        if (c->pows[h] > c->pows[imax])
            imax = h;
    const double pmax = c->pows[imax];
    double second = -HUGE_VAL;
    for (int h = 0; h < c->numharm; h++)
        if (h != imax && c->pows[h] > second)
            second = c->pows[h];
    if (c->numharm == 1)
        second = pmax;   /* one harmonic: its own power is also the second largest */
    if ((rules & ACCEL_SIFT_R_HARMPOWCUTOFF) && pmax < p->harm_pow_cutoff) {
        set_note(cw, "All harmonics have power < %g", p->harm_pow_cutoff);
        return ACCEL_SIFT_BAD_HARMPOWCUTOFF;
    }
    if (rules & ACCEL_SIFT_R_ROGUEHARMPOW) {
        /* one high-numbered harmonic far above all the others */
        if ((c->numharm >= 8 && imax > 4 && pmax > 2.0 * second) ||
            (c->numharm >= 4 && imax > 2 && pmax > 3.0 * second)) {
            set_note(cw, "High-numbered harmonic (%d) has too much power", imax);
            return ACCEL_SIFT_BAD_ROGUEHARMPOW;
        }
    }
    return 0;
}

int accel_sift_reject(accel_sift *s, const accel_sift_params *p, unsigned rules)
{
    clist *good = &s->list[ACCEL_SIFT_GOOD];
    int moved = 0;
    /* from the end of the list, so that each reject list receives its
     * candidates in the reverse of their good-list order */
    for (int i = good->n - 1; i >= 0; i--) {
        cand_t *c = good->v[i];
        const int to = failing_rule(c, c, p, rules);
        if (to) {
            move_good(s, i, to);
            moved++;
        }
    }
    return moved;
}

/* ---- duplicates across DM trials (python/presto/sifting.py:660-721) ---- */

int accel_sift_remove_duplicates(accel_sift *s, const accel_sift_params *p)
{
    clist *good = &s->list[ACCEL_SIFT_GOOD];
    stable_sort(good, by_r);
    int i = 0, moved = 0;
    while (i < good->n) {
        /* the run of candidates within r_err of the one at i (the list is
         * sorted by r, so the run is contiguous) */
        const double r0 = good->v[i]->r;
        int end = i + 1;
        while (end < good->n && fabs(good->v[end]->r - r0) < p->r_err)
            end++;
        if (end == i + 1) {
            i++;
            continue;
        }
        int best = i;
        for (int k = i + 1; k < end; k++)
            if (good->v[k]->sigma > good->v[best]->sigma)
                best = k;
        cand_t *b = good->v[best];
        /* the others become hits of the strongest, taken from the end of
         * the run; the strongest ends up at i and is compared again, as
         * its r may reach further */
        for (int k = end - 1; k >= i; k--) {
            if (k == best)
                continue;
            cand_t *m = good->v[k];
            if (add_hits(b, m->hits, m->nhits)) {
                accel_set_error("accel_sift_remove_duplicates: out of host memory");
                return -1;
            }
            set_note(m, "This candidate is a duplicate of %s:%d", b->file, b->candnum);
            move_good(s, k, ACCEL_SIFT_DUPLICATE);
            moved++;
        }
    }
    stable_sort(good, by_sigma_down);
    return moved;
}

/* ---- DM behaviour (python/presto/sifting.py:829-919) ---- */

static int cmp_double(const void *a, const void *b)
{
    const double x = *(const double *) a, y = *(const double *) b;
    return (x > y) - (x < y);
}

int accel_sift_remove_dm_problems(accel_sift *s, int min_num_dms, const double *dms, int ndms,
                                  double low_dm_cutoff)
{
    clist *good = &s->list[ACCEL_SIFT_GOOD];
    /* the DM grid: distinct values in ascending order, matched by their
     * two-decimal text */
    double *grid = (double *) malloc(sizeof(double) * (size_t) (ndms > 0 ? ndms : 1));
    char(*gtxt)[32] = (char(*)[32]) malloc(32 * (size_t) (ndms > 0 ? ndms : 1));
    if (!grid || !gtxt) {
        free(grid);
        free(gtxt);
        accel_set_error("accel_sift_remove_dm_problems: out of host memory");
        return -1;
    }
    memcpy(grid, dms, sizeof(double) * (size_t) ndms);
    qsort(grid, (size_t) ndms, sizeof(double), cmp_double);
    int ng = 0;
    for (int i = 0; i < ndms; i++)
        if (ng == 0 || grid[i] != grid[ng - 1])
            grid[ng++] = grid[i];
    for (int i = 0; i < ng; i++)
        snprintf(gtxt[i], 32, "%.2f", grid[i]);
    stable_sort(good, by_sigma_down);
    int moved = 0, rc = 0;
    for (int i = good->n - 1; i >= 0; i--) {
        cand_t *c = good->v[i];
        if (c->nhits < min_num_dms) {
            set_note(c, "Candidate has only %d DM hits. This is less than minimum for 'good' "
                        "cands (%d hits)", c->nhits, min_num_dms);
            move_good(s, i, ACCEL_SIFT_BAD_DMPROBLEM);
            moved++;
            continue;
        }
        int imax = 0;
        for (int h = 1; h < c->nhits; h++)
            if (c->hits[h].sigma > c->hits[imax].sigma)
                imax = h;
        if (c->hits[imax].dm <= low_dm_cutoff) {
            set_note(c, "Hit with max sigma (%g) has dm (%.2f) <= low DM cutoff (%.2f) ",
                     c->hits[imax].sigma, c->hits[imax].dm, low_dm_cutoff);
            move_good(s, i, ACCEL_SIFT_BAD_DMPROBLEM);
            moved++;
            continue;
        }
        if (c->nhits > 1) {
            /* the hits' DM-grid indices must include two neighbours */
            sort_hits_dm(c);
            int prev = -1, closest = 1 << 30;
            for (int h = 0; h < c->nhits; h++) {
                char t[32];
                snprintf(t, sizeof(t), "%.2f", c->hits[h].dm);
                int g = -1;
                for (int k = 0; k < ng; k++)
                    if (!strcmp(t, gtxt[k])) {
                        g = k;
                        break;
                    }
                if (g < 0) {
                    accel_set_error("accel_sift_remove_dm_problems: hit DM %s is not in the "
                                    "DM list", t);
                    rc = -1;
                    goto done;
                }
                if (h > 0 && g - prev < closest)
                    closest = g - prev;
                prev = g;
            }
            if (closest > 1) {
                set_note(c, "DM list of hits has gaps (i.e. consecutive DMs don't have hits).");
                move_good(s, i, ACCEL_SIFT_BAD_DMPROBLEM);
                moved++;
            }
        }
    }
done:
    free(grid);
    free(gtxt);
    return rc ? rc : moved;
}

/* ---- harmonics of stronger candidates (python/presto/sifting.py:723-827) ---- */

/* The ratios (numerator, denominator) besides whole harmonics that mark a
 * candidate as related to a stronger one. */
static const double OTHER_RATIOS[10][2] = {{3, 2}, {5, 2}, {2, 3}, {4, 3}, {5, 3},
                                          {3, 4}, {5, 4}, {2, 5}, {3, 5}, {4, 5}};

/* How candidate frequency fh relates to the stronger fund's ff within
 * f_err: writes the relation's name and returns 1, or 0. */
static int harmonic_relation(double ff, double fh, double f_err, char *name, size_t len)
{
    int found = 0;
    for (int k = 1; k <= 16 && !found; k++) {
        const double kk = k;
        if (fabs(ff - fh * kk) < f_err * kk) {
            snprintf(name, len, "1/%dth", k);       /* fh is a sub-harmonic */
            found = 1;
        } else if (fabs(ff - fh / kk) < f_err / kk) {
            snprintf(name, len, k == 2 ? "%dnd" : "%dth", k);   /* fh is a harmonic */
            found = 1;
        }
    }
    /* a fractional ratio, when one matches, names the relation */
    for (int i = 0; i < 10; i++) {
        const double fac = OTHER_RATIOS[i][0] / OTHER_RATIOS[i][1];
        if (fabs(ff - fh * fac) < f_err * fac) {
            snprintf(name, len, "%d/%dth", (int) OTHER_RATIOS[i][1], (int) OTHER_RATIOS[i][0]);
            return 1;
        }
    }
    return found;
}

int accel_sift_remove_harmonics(accel_sift *s, const accel_sift_params *p)
{
    clist *good = &s->list[ACCEL_SIFT_GOOD];
    stable_sort(good, by_sigma_down);
    if (good->n < 2)
        return 0;
    const double f_err = p->r_err / good->v[0]->T;
    int moved = 0;
    /* every candidate, strongest first, removes the weaker candidates that
     * are its harmonics, checked from the weakest up */
    for (int i = 0; i < good->n - 1; i++) {
        const cand_t *fund = good->v[i];
        const double ff = fund->f;
        for (int j = good->n - 1; j > i; j--) {
            cand_t *h = good->v[j];
            char rel[32];
            if (!harmonic_relation(ff, h->f, f_err, rel, sizeof(rel)))
                continue;
            set_note(h, "This candidate (P=%.4f s, DM=%.2f) is a harmonic (%s) of %s:%d "
                        "(P=%.4f s, DM=%.2f).",
                     h->p, h->dm, rel, fund->file, fund->candnum, fund->p, fund->dm);
            move_good(s, j, ACCEL_SIFT_BAD_HARMONIC);
            moved++;
        }
    }
    return moved;
}

int accel_sift_sort_sigma(accel_sift *s)
{
    stable_sort(&s->list[ACCEL_SIFT_GOOD], by_sigma_down);
    return 0;
}

/* Moves every candidate of `from` into s (emptying from): the good list,
 * the duplicates and each reject list are appended. */
int accel_sift_extend(accel_sift *s, accel_sift *from)
{
    for (int l = 0; l < ACCEL_SIFT_NLISTS; l++) {
        for (int i = 0; i < from->list[l].n; i++)
            if (cl_push(&s->list[l], from->list[l].v[i])) {
                accel_set_error("accel_sift_extend: out of host memory");
                return -1;
            }
        from->list[l].n = 0;
    }
    return 0;
}

/* ---- inspection ---- */

int accel_sift_count(const accel_sift *s, int list)
{
    if (!s || list < 0 || list >= ACCEL_SIFT_NLISTS)
        return -1;
    return s->list[list].n;
}

int accel_sift_get(const accel_sift *s, int list, int i, accel_sift_cand *out,
                   double *harm_pows, double *harm_phases, double *hits, int maxhits)
{
    if (!s || list < 0 || list >= ACCEL_SIFT_NLISTS || i < 0 || i >= s->list[list].n)
        return -1;
    const cand_t *c = s->list[list].v[i];
    if (out) {
        memset(out, 0, sizeof(*out));
        snprintf(out->path, sizeof(out->path), "%s", c->path);
        snprintf(out->dmstr, sizeof(out->dmstr), "%s", c->dmstr);
        snprintf(out->note, sizeof(out->note), "%s", c->note);
        out->candnum = c->candnum;
        out->numharm = c->numharm;
        out->sigma = c->sigma;
        out->ipow = c->ipow;
        out->cpow = c->cpow;
        out->r = c->r;
        out->z = c->z;
        out->T = c->T;
        out->snr = c->snr;
        out->nhits = c->nhits;
    }
    for (int h = 0; h < c->numharm; h++) {
        if (harm_pows)
            harm_pows[h] = c->pows[h];
        if (harm_phases)
            harm_phases[h] = c->phs[h];
    }
    for (int h = 0; h < c->nhits && h < maxhits; h++) {
        hits[3 * h] = c->hits[h].dm;
        hits[3 * h + 1] = c->hits[h].snr;
        hits[3 * h + 2] = c->hits[h].sigma;
    }
    return c->nhits;
}

/* ---- text outputs (python/presto/sifting.py:921-1071) ---- */

static FILE *open_out(const char *path)
{
    return path ? fopen(path, "w") : stdout;
}

static void close_out(FILE *f, const char *path)
{
    if (path)
        fclose(f);
    else
        fflush(f);
}

static void header(FILE *f)
{
    /* the column titles, each centred in its field */
    static const struct {
        const char *t;
        int w;
    } col[] = {{"file:candnum", 66}, {"DM", 9},   {"SNR", 8},     {"sigma", 8},
               {"numharm", 9},      {"ipow", 9}, {"cpow", 9},    {"P(ms)", 14},
               {"r", 12},           {"z", 8},    {"numhits", 9}};
    fputc('#', f);
    for (size_t i = 0; i < sizeof(col) / sizeof(col[0]); i++) {
        const int len = (int) strlen(col[i].t), pad = col[i].w - len;
        /* an odd pad puts its extra space on the left when the field
         * width is odd, else on the right (as Python's str.center) */
        const int left = pad / 2 + (pad & col[i].w & 1);
        fprintf(f, "%*s%s%*s", left, "", col[i].t, pad - left, "");
    }
    fputc('\n', f);
}

static void cand_line(FILE *f, const cand_t *c)
{
    char name[600];
    snprintf(name, sizeof(name), "%s:%d", c->file, c->candnum);
    fprintf(f, "%-65s   %7.2f  %6.2f  %6.2f    %2d     %7.1f  %7.1f  %12.6f  %10.2f  %8.2f  (%d)\n",
            name, c->dm, c->snr, c->sigma, c->numharm, c->ipow, c->cpow,
            1000.0 * c->p, c->r, c->z, c->nhits);
}

int accel_sift_write(const accel_sift *s, const char *path)
{
    FILE *f = open_out(path);
    if (!f) {
        accel_set_error("accel_sift_write: cannot open %s", path);
        return -1;
    }
    header(f);
    const clist *good = &s->list[ACCEL_SIFT_GOOD];
    for (int i = 0; i < good->n; i++) {
        cand_t *c = good->v[i];
        cand_line(f, c);
        if (c->nhits > 1) {
            sort_hits_dm(c);
            for (int h = 0; h < c->nhits; h++) {
                fprintf(f, "  DM=%6.2f SNR=%5.2f Sigma=%5.2f   ", c->hits[h].dm, c->hits[h].snr,
                        c->hits[h].sigma);
                for (int k = 0; k < (int) (c->hits[h].sigma / 3.0); k++)
                    fputc('*', f);
                fputc('\n', f);
            }
        }
    }
    close_out(f, path);
    return 0;
}

int accel_sift_write_report(const accel_sift *s, const char *path)
{
    FILE *f = open_out(path);
    if (!f) {
        accel_set_error("accel_sift_write_report: cannot open %s", path);
        return -1;
    }
    header(f);
    for (int l = ACCEL_SIFT_BAD_KNOWNBIRDS; l < ACCEL_SIFT_NLISTS; l++)
        for (int i = 0; i < s->list[l].n; i++) {
            cand_line(f, s->list[l].v[i]);
            fprintf(f, "    Note: %s\n\n", s->list[l].v[i]->note);
        }
    close_out(f, path);
    return 0;
}

int accel_sift_write_summary(const accel_sift *s, const char *path)
{
    FILE *f = open_out(path);
    if (!f) {
        accel_set_error("accel_sift_write_summary: cannot open %s", path);
        return -1;
    }
    static const struct {
        int list;
        const char *what;
    } rows[] = {{ACCEL_SIFT_BAD_KNOWNBIRDS, "Known RFI rejects:"},
                {ACCEL_SIFT_BAD_SHORTPERIOD, "Short period rejects:"},
                {ACCEL_SIFT_BAD_LONGPERIOD, "Long period rejects:"},
                {ACCEL_SIFT_BAD_THRESHOLD, "Missed threshold:"},
                {ACCEL_SIFT_BAD_HARMPOWCUTOFF, "No good harmonics:"},
                {ACCEL_SIFT_BAD_ROGUEHARMPOW, "One bad harmonic:"},
                {ACCEL_SIFT_DUPLICATE, "Duplicate candidates:"},
                {ACCEL_SIFT_BAD_HARMONIC, "Harmonic candidates:"},
                {ACCEL_SIFT_BAD_DMPROBLEM, "Candidates with DM problems:"}};
    fprintf(f, "   Candlist contains %d 'good' candidates\n", s->list[ACCEL_SIFT_GOOD].n);
    for (size_t i = 0; i < sizeof(rows) / sizeof(rows[0]); i++)
        fprintf(f, "      # %-28s %d\n", rows[i].what, s->list[rows[i].list].n);
    close_out(f, path);
    return 0;
}
