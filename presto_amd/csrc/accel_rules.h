/*
 * accel_rules.h -- the search's order- and rounding-sensitive host rules,
 * in one place for the search handle (accel_host.c) and the accel.h
 * surface (accel_surface.c), which must agree bit for bit:
 *
 *   - index rules, src/accel_utils.c:36, 54-102;
 *   - kernel half-widths and FFT lengths, src/responses.c:27-84,
 *     src/corr_prep.c:3-35, src/accel_utils.c:137-150;
 *   - the harmonic relations of eliminate_harmonics,
 *     src/accel_utils.c:472-549;
 *   - insert_new_accelcand, src/accel_utils.c:387-469, as a macro that
 *     instantiates the one merge over a given list type.
 */
#ifndef ACCEL_RULES_H
#define ACCEL_RULES_H
#include <math.h>
#include "presto_accel.h"

/* include/accel.h:18-35, include/presto.h:99-107 */
#define ACCEL_NUMBETWEEN 2
#define ACCEL_DR 0.5
#define ACCEL_RDR 2
#define ACCEL_DZ 2
#define ACCEL_RDZ 0.5
#define ACCEL_DW 20
#define ACCEL_RDW 0.05
#define ACCEL_CLOSEST_R 15.0
#define NUMFINTBINS 16
#define DBLCORRECT 1e-14

static inline int ar_nearest_int(double x) { return (int) (x < 0 ? x - 0.5 : x + 0.5); }
static inline double ar_required_r(double hf, double rfull)
{
    return rint(ACCEL_RDR * rfull * hf) * ACCEL_DR;
}
static inline int ar_required_z(double hf, double zfull)
{
    return ar_nearest_int(ACCEL_RDZ * zfull * hf) * ACCEL_DZ;
}
static inline int ar_required_w(double hf, double wfull)
{
    return ar_nearest_int(ACCEL_RDW * wfull * hf) * ACCEL_DW;
}

static inline int ar_z_halfwidth(double z) { return (int) (0.5 * 1.1 * fabs(z)) + NUMFINTBINS; }

static inline int ar_w_halfwidth(double z, double w)
{
    if (fabs(w) < 1.0e-7)
        return ar_z_halfwidth(z);
    double r0 = 0.5 * (w / 6.0 - z), r1 = 0.5 * (w / 6.0 + z);
    double maxdev = fabs(r0) > fabs(r1) ? fabs(r0) : fabs(r1);
    double u_ext = 0.5 - z / w;
    if (u_ext > 0.0 && u_ext < 1.0) {
        double r_ext = 0.5 * w * u_ext * u_ext + (z - w / 2.0) * u_ext + r0;
        if (fabs(r_ext) > maxdev)
            maxdev = fabs(r_ext);
    }
    return (int) (1.1 * maxdev) + NUMFINTBINS;
}

static inline long long ar_next2(long long x)
{
    long long t = 1;
    while (t < x)
        t <<= 1;
    return t;
}

static inline int ar_next_good_fftlen(int n)
{
    static const int lens[17] = {128, 192, 256, 384, 512, 768, 1024, 1280, 2048,
                                 4096, 5120, 7680, 10240, 12288, 15360, 16384, 25600};
    if (n <= lens[0])
        return lens[0];
    if (n > lens[16])
        return (int) ar_next2(n);
    int i = 0;
    while (n > lens[i])
        i++;
    return lens[i];
}

static inline int ar_calc_fftlen(int numharm, int harmnum, int zmax, int wmax, int corr_uselen)
{
    double hf = (double) harmnum / (double) numharm;
    int bins_needed = (int) ceil(corr_uselen * hf) + 2;
    int end_effects = 2 * ACCEL_NUMBETWEEN *
        ar_w_halfwidth(ar_required_z(hf, zmax), ar_required_w(hf, wmax));
    return ar_next_good_fftlen(bins_needed + end_effects);
}

/* eliminate_harmonics' test: is other's r a harmonic or simple ratio of
 * cur's (within 1.5 bins)? */
static inline int ar_harmonically_related(double cr, double orr)
{
    static const double ratios[23][2] = {
        {3, 2}, {5, 2}, {2, 3}, {4, 3}, {5, 3}, {3, 4}, {5, 4}, {2, 5},
        {3, 5}, {4, 5}, {5, 6}, {2, 7}, {3, 7}, {4, 7}, {3, 8}, {5, 8},
        {2, 9}, {3, 10}, {2, 11}, {3, 11}, {2, 13}, {3, 13}, {2, 15}};
    const double tooclose = 1.5;
    for (int ii = 1; ii <= 16; ii++)
        if (fabs(cr / ii - orr) < tooclose)
            return 1;
    for (int ii = 1; ii <= 16; ii++)
        if (fabs(cr * ii - orr) < tooclose)
            return 1;
    for (int k = 0; k < 23; k++)
        if (fabs(cr * ratios[k][0] / ratios[k][1] - orr) < tooclose)
            return 1;
    return 0;
}

/* insert_new_accelcand over a singly linked list of node type NODE whose
 * successor is node->next.  The list is kept sorted by r; a new candidate
 * within ACCEL_CLOSEST_R bins of a neighbour replaces it only with a
 * higher sigma, and can replace both neighbours (the second is dropped).
 * Accessors: R(n), SIGMA(n); SET(n, p, s, h, r, z, w) overwrites a node;
 * NEW(p, s, h, r, z, w) makes a node; DROP(n) frees one. */
#define ACCEL_DEFINE_INSERT(NAME, NODE, R, SIGMA, SET, NEW, DROP)                       \
    static NODE *NAME(NODE *list, float power, float sigma, int numharm, double rr,      \
                      double zz, double ww, int *added)                                 \
    {                                                                                    \
        NODE *tmp = list, *prev = NULL;                                                  \
        double prev_diff = ACCEL_CLOSEST_R + 1.0, next_diff;                             \
        *added = 0;                                                                      \
        if (!list) {                                                                     \
            *added = 1;                                                                  \
            return NEW(power, sigma, numharm, rr, zz, ww);                               \
        }                                                                                \
        while (tmp->next && R(tmp) < rr) {                                               \
            prev = tmp;                                                                  \
            tmp = tmp->next;                                                             \
        }                                                                                \
        next_diff = fabs(rr - R(tmp));                                                   \
        if (prev)                                                                        \
            prev_diff = fabs(rr - R(prev));                                              \
        if (prev_diff < ACCEL_CLOSEST_R) {                                               \
            if (SIGMA(prev) < sigma) {                                                   \
                SET(prev, power, sigma, numharm, rr, zz, ww);                            \
                *added = 1;                                                              \
            }                                                                            \
            if (next_diff < ACCEL_CLOSEST_R && SIGMA(tmp) < sigma) {                     \
                if (*added) {                                                            \
                    prev->next = tmp->next;                                              \
                    DROP(tmp);                                                           \
                } else {                                                                 \
                    SET(tmp, power, sigma, numharm, rr, zz, ww);                         \
                    *added = 1;                                                          \
                }                                                                        \
            }                                                                            \
        } else if (next_diff < ACCEL_CLOSEST_R) {                                        \
            if (SIGMA(tmp) < sigma) {                                                    \
                SET(tmp, power, sigma, numharm, rr, zz, ww);                             \
                *added = 1;                                                              \
            }                                                                            \
        } else {                                                                         \
            NODE *n_ = NEW(power, sigma, numharm, rr, zz, ww);                           \
            *added = 1;                                                                  \
            if (!tmp->next && R(tmp) < (rr - ACCEL_CLOSEST_R)) {                         \
                tmp->next = n_;                                                          \
                return list;                                                             \
            }                                                                            \
            if (prev) {                                                                  \
                prev->next = n_;                                                         \
                n_->next = tmp;                                                          \
            } else {                                                                     \
                n_->next = list;                                                         \
                return n_;                                                               \
            }                                                                            \
        }                                                                                \
        return list;                                                                     \
    }

/* eliminate_harmonics over the same kind of list: every later candidate
 * harmonically related to an earlier one is dropped (DROP(n)) */
#define ACCEL_DEFINE_ELIMINATE(NAME, NODE, R, DROP)                                      \
    static NODE *NAME(NODE *cands, int *numcands)                                        \
    {                                                                                    \
        NODE *cur = cands;                                                               \
        if (!cur)                                                                        \
            return cands;                                                                \
        while (cur->next) {                                                              \
            NODE *before = cur, *other = cur->next;                                      \
            do {                                                                         \
                if (ar_harmonically_related(R(cur), R(other))) {                         \
                    NODE *gone = other;                                                  \
                    other = other->next;                                                 \
                    before->next = other;                                                \
                    DROP(gone);                                                          \
                    (*numcands)--;                                                       \
                } else {                                                                 \
                    before = other;                                                      \
                    other = other->next;                                                 \
                }                                                                        \
            } while (other);                                                             \
            if (cur->next)                                                               \
                cur = cur->next;                                                         \
        }                                                                                \
        return cands;                                                                    \
    }

#endif
