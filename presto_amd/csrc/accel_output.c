/*
 * accel_output.c -- candidate properties and the accelsearch output files.
 *
 *   accel_calc_props / accel_calc_rzwerrs   src/characteristics.c:195-262, 330-396
 *   accel_nice_output                       src/output.c:6-216 (value(error) text)
 *   accel_search_write_outputs              src/accel_utils.c:629-979 (ACCEL text:
 *                                           fundamentals table, harmonics table,
 *                                           .inf appended) and src/accelsearch.c:
 *                                           281-284 (.cand = fourierprops[])
 *
 * The text layout (column titles, widths, centring, the value(error)
 * notation) is a file format other PRESTO tools parse (sifting.py reads
 * these files), so it is kept exactly; the code that produces it is
 * organised around column tables.  The pulsar-catalogue comparison of the
 * Notes column (comp_psr_to_cand) is not built: Notes carry only the
 * candidate-to-candidate comparison, as for data whose .inf telescope is
 * "None".
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "presto_accel.h"
#include "accel_internal.h"

#define PI 3.14159265358979323846
#define TWOPI 6.28318530717958647692
#define SOL 299792458.0
#define DBLCORRECT 1e-14

int accel_calc_props(const accel_rderivs *d, double r, double z, double w,
                     accel_fourierprops *p)
{
    if (fabs(d->pow) < DBLCORRECT || d->locpow < DBLCORRECT)
        return -1;
    /* the record is written whole to .cand files: clear its padding too */
    memset(p, 0, sizeof(*p));
    const double pw = d->pow / d->locpow, d2 = d->d2pow / d->locpow;
    p->pur = sqrt(1.5 * fabs(d2) / pw) / PI;
    if (fabs(pw) < DBLCORRECT)
        p->pur = 1.0;
    const double pur = p->pur;   /* the stored (FP32) purity feeds the errors */
    p->purerr = 1.0 / (pur * sqrt(10.0 * pw));
    p->r = r;
    p->rerr = 3.0 / (PI * pur * sqrt(6.0 * pw));
    p->z = z;
    p->zerr = 3.0 * sqrt(10.0) / (PI * pur * pur * sqrt(pw));
    p->w = w;
    p->werr = 6.0 * sqrt(105.0) / (PI * pur * pur * pur * sqrt(pw));
    p->pow = pw;
    p->powerr = sqrt(2.0 * pw);
    p->sig = sqrt(2.0 * pw - log(PI * pw));
    p->rawpow = d->pow;
    p->phs = d->phs;
    p->phserr = 1.0 / p->powerr;
    p->cen = -d->dphs / TWOPI;
    p->cenerr = 1.0 / sqrt(24.0 * pw);
    p->locpow = d->locpow;
    return 0;
}

int accel_calc_rzwerrs(const accel_fourierprops *p, double T, accel_rzwerrs *e)
{
    if (T <= 0.0)
        return -1;
    const double T2 = T * T, T3 = T2 * T;
    const double r2 = p->r * p->r, r4 = r2 * r2;
    const double sr2 = (double) (p->rerr * p->rerr), sz2 = (double) (p->zerr * p->zerr);
    const double z2 = p->z * p->z;
    const double tmp = p->r * p->w - 3 * z2;
    e->f = p->r / T;
    e->ferr = p->rerr / T;
    e->fd = p->z / T2;
    e->fderr = p->zerr / T2;
    e->fdd = p->w / T3;
    e->fdderr = p->werr / T3;
    e->p = T / p->r;
    e->perr = T * p->rerr / r2;
    e->pd = -p->z / r2;
    e->pderr = sqrt(4 * z2 * sr2 / (r4 * r2) + sz2 / r4);
    e->pdd = (p->w == 0.0) ? 0.0 : (2 * z2 - p->r * p->w) / (T * r2 * p->r);
    e->pdderr = sqrt(((double) (p->werr * p->werr) * r4 + 16 * sz2 * r2 + 4 * tmp * tmp * sr2) /
                     (r4 * r4 * T2));
    return 0;
}

static double signed_like(double a, double b)
{
    return b < 0.0 ? -fabs(a) : fabs(a);
}

/* value(error) with a one-digit error, src/output.c:6-113 */
static int nice1(char *out, double val, double err, int len)
{
    char t[64];
    int errexp = (fabs(err) == 0.0) ? 0 : (int) floor(log10(fabs(err)));
    int errval = (int) floor(fabs(err) * pow(10.0, -errexp) + DBLCORRECT + 0.5);
    if (errval == 10) {
        errval = 1;
        errexp++;
    }
    double rndval = pow(10.0, errexp) * floor(val * pow(10.0, -errexp) + 0.5);
    int nint;
    if (fabs(val) == 0.0) {
        nint = 1;
    } else {
        nint = (int) ceil(log10(fabs(val)));
        if (nint == 0)
            nint++;
    }
    int nfrac = -errexp;
    int totprec = nint + nfrac;
    int outexp = (fabs(rndval) == 0.0) ? 0 : (int) floor(log10(fabs(rndval)));
    double outmant = rndval * pow(10.0, -outexp);
    if (fabs(1.0 - outmant) < DBLCORRECT || fabs(-1.0 - outmant) < DBLCORRECT)
        totprec++;
    if (outexp >= 0 && errexp > 0 && outexp > errexp)
        snprintf(t, sizeof(t), "% .*f(%d)x10^%d", totprec - 1, signed_like(outmant, rndval),
                 errval, outexp);
    else if (outexp >= 0 && errexp > 0 && outexp == errexp)
        snprintf(t, sizeof(t), "% d(%d)x10^%d", (int) signed_like(outmant, rndval), errval,
                 outexp);
    else if (outexp < -4 && outexp >= errexp)
        snprintf(t, sizeof(t), "% .*f(%d)x10^%d", totprec - 1, signed_like(outmant, rndval),
                 errval, outexp);
    else if (outexp < errexp && errexp != 0)
        snprintf(t, sizeof(t), "% d(%d)x10^%d",
                 (int) (signed_like(outmant, rndval) + DBLCORRECT), errval, errexp);
    else if (nfrac == 0 && fabs(rndval) < 1.0e-15)
        snprintf(t, sizeof(t), "% d(%d)", (int) fabs(rndval), errval);
    else if (fabs(rndval) <= DBLCORRECT && errexp < -5)
        snprintf(t, sizeof(t), "0.0(%d)x10^%d", errval, errexp + 1);
    else
        snprintf(t, sizeof(t), "% .*f(%d)", nfrac, rndval, errval);
    if (len == 0)
        return snprintf(out, 80, "%s", t);
    return snprintf(out, 80, "%*s", len, t);
}

/* value(error) with a two-digit error, src/output.c:114-216 */
static int nice2(char *out, double val, double err, int len)
{
    char t[64];
    int errexp = (fabs(err) == 0.0) ? 0 : (int) floor(log10(fabs(err)));
    int errval = (int) floor(fabs(err) * pow(10.0, -errexp + 1) + DBLCORRECT + 0.5);
    if (errval == 100) {
        errval = 10;
        errexp++;
    }
    double rndval = pow(10.0, errexp - 1) * floor(val * pow(10.0, -errexp + 1) + 0.5);
    int outexp = (rndval == 0.0) ? 0 : (int) floor(log10(fabs(rndval)));
    int valexp = (val == 0.0) ? 0 : (int) floor(log10(fabs(val)));
    double outmant = rndval * pow(10.0, -outexp);
    if (outexp < -4) {
        int d = outexp - errexp + 1;
        if (d < 2) {
            if (d == 0)
                snprintf(t, sizeof(t), "% .1f(%d.%d)x10^%d", outmant / 10, errval / 10,
                         errval % 10, outexp + 1);
            else
                snprintf(t, sizeof(t), "% .1f(%d.%d)x10^%d", outmant, errval / 10,
                         errval % 10, outexp);
        } else {
            snprintf(t, sizeof(t), "% .*f(%.2d)x10^%d", d, outmant, errval, outexp);
        }
    } else if (errexp - valexp > 1 && (errexp != 0 && errexp != 1)) {
        snprintf(t, sizeof(t), "0.0(%d.%d)x10^%d", errval / 10, errval % 10, errexp);
    } else if (errexp == 0) {
        if (fabs(rndval) < 0.1)
            rndval = fabs(rndval);
        snprintf(t, sizeof(t), "% .1f(%d.%d)", rndval, errval / 10, errval % 10);
    } else if (errexp < 0) {
        if (fabs(rndval) * pow(10.0, -errexp + 1) < 0.1)
            rndval = fabs(rndval);
        if (fabs(rndval) <= DBLCORRECT && errexp < -4)
            snprintf(t, sizeof(t), "0.0(%d.%d)x10^%d", errval / 10, errval % 10, errexp);
        else
            snprintf(t, sizeof(t), "% .*f(%.2d)", -errexp + 1, rndval, errval);
    } else if (errexp - outexp == 0) {
        if (errexp == 1)
            snprintf(t, sizeof(t), "% d(%.2d)", (int) rndval, errval);
        else
            snprintf(t, sizeof(t), "% .1f(%d.%d)x10^%d", outmant, errval / 10, errval % 10,
                     errexp);
    } else if (errexp - outexp == 1) {
        if (fabs(outmant) < 0.99999999999999)
            outmant = fabs(outmant);
        snprintf(t, sizeof(t), "% .1f(%d.%d)x10^%d", outmant / 10.0, errval / 10, errval % 10,
                 errexp);
    } else if ((errexp > 1 && outexp > 1) && errexp - outexp <= -1) {
        snprintf(t, sizeof(t), "% .*f(%.2d)x10^%d", outexp - errexp + 1, outmant, errval,
                 outexp);
    } else if (errexp == 1 && outexp > 0) {
        snprintf(t, sizeof(t), "% d(%.2d)", (int) rndval, errval);
    } else {
        return -1;   /* the reference's "undefined condition" */
    }
    if (len == 0)
        return snprintf(out, 80, "%s", t);
    return snprintf(out, 80, "%*s", len, t);
}

int accel_nice_output(char *out, double val, double err, int len, int numerr)
{
    if (numerr == 1)
        return nice1(out, val, err, len);
    if (numerr == 2)
        return nice2(out, val, err, len);
    return -1;
}

/* src/accel_utils.c:629-646: s centred in a field of width characters
 * (odd spare space goes to the right), truncated if longer */
static void centre(char *out, const char *s, int width)
{
    int len = (int) strlen(s);
    memset(out, ' ', width);
    out[width] = '\0';
    if (len >= width)
        memcpy(out, s, width);
    else
        memcpy(out + (width - len) / 2, s, len);
}

static void put_cell(FILE *f, const char *s, int width, const char *sep)
{
    char c[128];
    centre(c, s, width);
    fprintf(f, "%s%s", c, sep);
}

static void put_valerr(FILE *f, double val, double err, int numerr, int width)
{
    char t[96];
    if (accel_nice_output(t, val, err, 0, numerr) < 0)
        snprintf(t, sizeof(t), "%g(%g)", val, err);
    put_cell(f, t, width, "  ");
}

/* src/accel_utils.c:105-134 */
static void compare_cands(const accel_fourierprops *l, int n, char *notes)
{
    static const char blank[21] = "                    ";
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < n; j++) {
            if (i == j)
                continue;
            char t[32];
            if (fabs(l[i].r - l[j].r) < 15.0 && fabs(l[i].z - l[j].z) > 1.0 &&
                l[i].pow > l[j].pow) {
                if (strncmp(notes + j * 20, blank, 20) == 0) {
                    snprintf(t, sizeof(t), "SL? of Cand %d", i + 1);
                    memset(notes + j * 20, 0, 20);
                    memcpy(notes + j * 20, t, strlen(t) < 20 ? strlen(t) : 20);
                }
                continue;
            }
            for (int k = 1; k < 61; k++) {
                if (fabs(l[i].r - l[j].r / k) < l[j].rerr * 3 &&
                    fabs(l[i].z - l[j].z / k) < l[j].zerr * 2) {
                    if (strncmp(notes + j * 20, blank, 20) == 0) {
                        snprintf(t, sizeof(t), "H %d of Cand %d", k, i + 1);
                        memset(notes + j * 20, 0, 20);
                        memcpy(notes + j * 20, t, strlen(t) < 20 ? strlen(t) : 20);
                        break;
                    }
                }
            }
        }
    }
}

typedef struct {
    const char *t1, *t2;
    int width, numerr;   /* numerr 0: not a value(error) column */
    int jerk;            /* printed only for jerk searches */
} col;

/* output_fundamentals columns, src/accel_utils.c:686-695 */
static const col FUND[13] = {
    {"", "Cand", 4, 0, 0},           {"", "Sigma", 5, 0, 0},
    {"Summed", "Power", 6, 0, 0},    {"Coherent", "Power", 8, 0, 0},
    {"Num", "Harm", 4, 0, 0},        {"Period", "(ms)", 16, 1, 0},
    {"Frequency", "(Hz)", 15, 1, 0}, {"FFT 'r'", "(bin)", 15, 2, 0},
    {"Freq Deriv", "(Hz/s)", 15, 1, 0}, {"FFT 'z'", "(bins)", 11, 2, 0},
    {"FFT 'w'", "(bins)", 11, 2, 1}, {"Accel", "(m/s^2)", 15, 2, 0},
    {"", "Notes", 20, 0, 0}};

/* output_harmonics columns, src/accel_utils.c:862-873 */
static const col HARM[15] = {
    {"", "Cand", 5, 0, 0},           {"", "Harm", 4, 0, 0},
    {"", "Sigma", 5, 0, 0},          {"Power /", "Loc Pow", 15, 2, 0},
    {"Raw", "Power", 11, 0, 0},      {"FFT 'r'", "(bin)", 18, 2, 0},
    {"Pred 'r'", "(bin)", 13, 0, 0}, {"FFT 'z'", "(bins)", 12, 2, 0},
    {"Pred 'z'", "(bins)", 9, 0, 0}, {"FFT 'w'", "(bins)", 12, 2, 1},
    {"Pred 'w'", "(bins)", 9, 0, 1}, {"Phase", "(rad)", 12, 2, 0},
    {"Centroid", "(0-1)", 10, 2, 0}, {"Purity", "<p> = 1", 10, 2, 0},
    {"", "Notes", 20, 0, 0}};

static void header(FILE *f, const col *c, int n, int jerk)
{
    for (int row = 0; row < 3; row++) {
        for (int i = 0; i < n; i++) {
            if (c[i].jerk && !jerk)
                continue;
            const int last = (i == n - 1);
            if (row < 2) {
                put_cell(f, row == 0 ? c[i].t1 : c[i].t2, c[i].width, last ? "\n" : "  ");
            } else {
                char d[64];
                memset(d, '-', c[i].width);
                d[c[i].width] = '\0';
                fprintf(f, "%s%s", d, last ? "\n" : "--");
            }
        }
    }
}

int accel_out_fundamentals(FILE *f, const accel_out_cand *cands,
                           const accel_fourierprops *props, int n, double T, int jerk)
{
    char *notes = (char *) malloc((size_t) 20 * (n ? n : 1) + 1);
    if (!notes)
        return -1;
    memset(notes, ' ', (size_t) 20 * n);
    compare_cands(props, n, notes);
    header(f, FUND, 13, jerk);
    for (int i = 0; i < n; i++) {
        const accel_out_cand *c = &cands[i];
        accel_rzwerrs e;
        accel_calc_rzwerrs(&props[i], T, &e);
        /* coherently summed power, src/accel_utils.c:780-797 */
        double cr = 0.0, ci = 0.0, phs0 = c->derivs[0].phs;
        for (int j = 0; j < c->numharm; j++) {
            const accel_rderivs *d = &c->derivs[j];
            double amp = sqrt(d->pow / d->locpow);
            double corr = phs0 - fmod((j + 1.0) * phs0, TWOPI);
            cr += amp * cos(d->phs + corr);
            ci += amp * sin(d->phs + corr);
        }
        char t[64];
        snprintf(t, sizeof(t), "%-4d", i + 1);
        put_cell(f, t, FUND[0].width, "  ");
        snprintf(t, sizeof(t), "%.2f", c->sigma);
        put_cell(f, t, FUND[1].width, "  ");
        snprintf(t, sizeof(t), "%.2f", c->power);
        put_cell(f, t, FUND[2].width, "  ");
        snprintf(t, sizeof(t), "%.2f", cr * cr + ci * ci);
        put_cell(f, t, FUND[3].width, "  ");
        snprintf(t, sizeof(t), "%d", c->numharm);
        put_cell(f, t, FUND[4].width, "  ");
        put_valerr(f, e.p * 1000.0, e.perr * 1000.0, FUND[5].numerr, FUND[5].width);
        put_valerr(f, e.f, e.ferr, FUND[6].numerr, FUND[6].width);
        put_valerr(f, props[i].r, props[i].rerr, FUND[7].numerr, FUND[7].width);
        put_valerr(f, e.fd, e.fderr, FUND[8].numerr, FUND[8].width);
        put_valerr(f, props[i].z, props[i].zerr, FUND[9].numerr, FUND[9].width);
        if (jerk)
            put_valerr(f, props[i].w, props[i].werr, FUND[10].numerr, FUND[10].width);
        double accel = props[i].z * SOL / (T * T * e.f);
        double accelerr = props[i].zerr * SOL / (T * T * e.f);
        put_valerr(f, accel, accelerr, FUND[11].numerr, FUND[11].width);
        fprintf(f, "  %.20s\n", notes + i * 20);
    }
    fprintf(f, "\n\n");
    free(notes);
    return 0;
}

int accel_out_harmonics(FILE *f, const accel_out_cand *cands, int n, double T, int jerk)
{
    header(f, HARM, 15, jerk);
    for (int i = 0; i < n; i++) {
        const accel_out_cand *c = &cands[i];
        for (int j = 0; j < c->numharm; j++) {
            accel_fourierprops hp;
            accel_rzwerrs e;
            if (accel_calc_props(&c->derivs[j], c->hirs[j], c->hizs[j], c->hiws[j], &hp))
                memset(&hp, 0, sizeof(hp));
            accel_calc_rzwerrs(&hp, T, &e);
            char t[64];
            if (j == 0)
                snprintf(t, sizeof(t), " %-4d", i + 1);
            else
                snprintf(t, sizeof(t), "     ");
            put_cell(f, t, HARM[0].width, "  ");
            snprintf(t, sizeof(t), "%-4d", j + 1);
            put_cell(f, t, HARM[1].width, "  ");
            snprintf(t, sizeof(t), "%.2f", accel_candidate_sigma(hp.pow, 1, 1));
            put_cell(f, t, HARM[2].width, "  ");
            put_valerr(f, hp.pow, hp.powerr, HARM[3].numerr, HARM[3].width);
            snprintf(t, sizeof(t), "%.3g", hp.rawpow);
            put_cell(f, t, HARM[4].width, "  ");
            put_valerr(f, hp.r, hp.rerr, HARM[5].numerr, HARM[5].width);
            snprintf(t, sizeof(t), "%.2f", c->r * (j + 1));
            put_cell(f, t, HARM[6].width, "  ");
            put_valerr(f, hp.z, hp.zerr, HARM[7].numerr, HARM[7].width);
            snprintf(t, sizeof(t), "%.2f", c->z * (j + 1));
            put_cell(f, t, HARM[8].width, "  ");
            if (jerk) {
                put_valerr(f, hp.w, hp.werr, HARM[9].numerr, HARM[9].width);
                snprintf(t, sizeof(t), "%.2f", c->w * (j + 1));
                put_cell(f, t, HARM[10].width, "  ");
            }
            put_valerr(f, hp.phs, hp.phserr, HARM[11].numerr, HARM[11].width);
            put_valerr(f, hp.cen, hp.cenerr, HARM[12].numerr, HARM[12].width);
            put_valerr(f, hp.pur, hp.purerr, HARM[13].numerr, HARM[13].width);
            fprintf(f, "  %.20s\n", "                    ");
        }
    }
    fprintf(f, "\n\n");
    return 0;
}

void accel_out_append_file(FILE *f, const char *path)
{
    FILE *in = path ? fopen(path, "r") : NULL;
    if (!in)
        return;
    char buf[4096];
    size_t k;
    while ((k = fread(buf, 1, sizeof(buf), in)) > 0)
        fwrite(buf, 1, k, f);
    fclose(in);
}

int accel_search_write_outputs(const accel_search *s, const char *accelnm, const char *candnm,
                               const char *inf_path)
{
    const accel_geom *g = accel_search_geom(s);
    const int n = accel_search_ncands(s);
    const int jerk = g->numw != 0;
    const int m = n ? n : 1;
    accel_cand *cands = (accel_cand *) malloc(sizeof(accel_cand) * m);
    accel_fourierprops *props = (accel_fourierprops *) calloc(m, sizeof(accel_fourierprops));
    accel_out_cand *oc = (accel_out_cand *) calloc(m, sizeof(accel_out_cand));
    accel_rderivs *der = (accel_rderivs *) calloc((size_t) m * ACCEL_MAXHARM,
                                                  sizeof(accel_rderivs));
    double *hi = (double *) calloc((size_t) 3 * m * ACCEL_MAXHARM, sizeof(double));
    FILE *f = NULL;
    int rc = -1;
    if (!cands || !props || !oc || !der || !hi)
        goto done;
    accel_search_cands(s, cands, n);
    if (n && accel_search_fund_props(s, props, n) != n)
        goto done;
    for (int i = 0; i < n; i++) {
        double *hr = hi + (size_t) 3 * i * ACCEL_MAXHARM;
        accel_rderivs *d = der + (size_t) i * ACCEL_MAXHARM;
        accel_search_cand_harmonics(s, i, hr, hr + ACCEL_MAXHARM, hr + 2 * ACCEL_MAXHARM, NULL,
                                    d, ACCEL_MAXHARM);
        oc[i].power = cands[i].power;
        oc[i].sigma = cands[i].sigma;
        oc[i].numharm = cands[i].numharm;
        oc[i].r = cands[i].r;
        oc[i].z = cands[i].z;
        oc[i].w = cands[i].w;
        oc[i].derivs = d;
        oc[i].hirs = hr;
        oc[i].hizs = hr + ACCEL_MAXHARM;
        oc[i].hiws = hr + 2 * ACCEL_MAXHARM;
    }
    f = fopen(accelnm, "w");
    if (!f)
        goto done;
    /* src/accel_utils.c:666-979 */
    accel_out_fundamentals(f, oc, props, n, g->T, jerk);
    accel_out_harmonics(f, oc, n, g->T, jerk);
    accel_out_append_file(f, inf_path);
    if (fclose(f)) {
        f = NULL;
        goto done;
    }
    f = NULL;
    if (candnm) {
        FILE *cf = fopen(candnm, "wb");
        if (!cf)
            goto done;
        if (n && fwrite(props, sizeof(accel_fourierprops), n, cf) != (size_t) n) {
            fclose(cf);
            goto done;
        }
        if (fclose(cf))
            goto done;
    }
    rc = 0;
done:
    if (f)
        fclose(f);
    free(cands);
    free(props);
    free(oc);
    free(der);
    free(hi);
    return rc;
}
