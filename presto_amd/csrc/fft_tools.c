/*
 * fft_tools.c -- PRESTO's standalone .fft producers on the MI355X path,
 * built three times: -DTOOL_REALFFT (realfft, src/realfft.c),
 * -DTOOL_REDNOISE (rednoise, src/rednoise.c) and -DTOOL_ZAPBIRDS (the
 * automatic mode of zapbirds, src/zapbirds.c:273-303).  Options, file
 * naming and messages follow clig/{realfft,rednoise,zapbirds}_cmd.cli.
 *
 * The transforms and the de-reddening run on the GPU (accel_realfft_forward,
 * accel_realfft_inverse, accel_deredden_gpu); time series lengths must be
 * even, as the reference's realfft requires.  -disk, -mem and -tmpdir are
 * accepted and have no effect
 * (every transform is in-core, in device memory); zapbirds' interactive
 * mode (-in / -out without -zap) is not built.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include "accel_surface.h"

static void *read_all(const char *path, size_t elsize, long long *count)
{
    FILE *f = fopen(path, "rb");
    if (!f)
        return NULL;
    fseek(f, 0, SEEK_END);
    long long bytes = ftell(f);
    fseek(f, 0, SEEK_SET);
    *count = bytes / (long long) elsize;
    void *buf = malloc((size_t) (*count ? *count : 1) * elsize);
    if (buf && *count && fread(buf, elsize, (size_t) *count, f) != (size_t) *count) {
        free(buf);
        buf = NULL;
    }
    fclose(f);
    return buf;
}

static int write_all(const char *path, const void *buf, size_t elsize, long long count)
{
    FILE *f = fopen(path, "wb");
    if (!f)
        return -1;
    size_t w = count ? fwrite(buf, elsize, (size_t) count, f) : 0;
    int rc = fclose(f);
    return (w == (size_t) count && rc == 0) ? 0 : -1;
}

/* root = everything before the last '.'; returns the suffix or NULL */
static const char *split(const char *in, char *root, size_t n)
{
    const char *dot = strrchr(in, '.');
    const char *slash = strrchr(in, '/');
    if (!dot || (slash && dot < slash)) {
        snprintf(root, n, "%s", in);
        return NULL;
    }
    snprintf(root, n, "%.*s", (int) (dot - in), in);
    return dot + 1;
}

/* src/fastffts.c:173-176: realfft takes even lengths */
static int even_len(long long n) { return n >= 4 && !(n & 1); }

#if defined(TOOL_REALFFT)

static void usage(void)
{
    printf("\nUsage:  realfft [-fwd] [-inv] [-del] [-disk] [-mem] [-rednoise]\n"
           "        [-startwidth startwidth] [-endwidth endwidth] [-endfreq endfreq]\n"
           "        [-tmpdir tmpdir] [-outdir outdir] [-device d] infiles ...\n\n");
}

int main(int argc, char **argv)
{
    int fwd = 0, inv = 0, del = 0, red = 0, startwidth = 6, endwidth = 100, device = 0;
    double endfreq = 6.0;
    const char *outdir = NULL;
    int first = argc;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (!strcmp(a, "-fwd")) fwd = 1;
        else if (!strcmp(a, "-inv")) inv = 1;
        else if (!strcmp(a, "-del")) del = 1;
        else if (!strcmp(a, "-disk") || !strcmp(a, "-mem")) ;
        else if (!strcmp(a, "-rednoise")) red = 1;
        else if (!strcmp(a, "-startwidth") && i + 1 < argc) startwidth = atoi(argv[++i]);
        else if (!strcmp(a, "-endwidth") && i + 1 < argc) endwidth = atoi(argv[++i]);
        else if (!strcmp(a, "-endfreq") && i + 1 < argc) endfreq = atof(argv[++i]);
        else if (!strcmp(a, "-tmpdir") && i + 1 < argc) ++i;
        else if (!strcmp(a, "-outdir") && i + 1 < argc) outdir = argv[++i];
        else if (!strcmp(a, "-device") && i + 1 < argc) device = atoi(argv[++i]);
        else if (a[0] == '-') {
            fprintf(stderr, "realfft: unknown option '%s'\n", a);
            usage();
            return 1;
        } else {
            first = i;
            break;
        }
    }
    if (first >= argc) {
        usage();
        return 1;
    }
    /* clig/realfft_cmd.cli ranges */
    if (startwidth < 2 || startwidth > 50 || endwidth < 50 || endwidth > 500 ||
        endfreq < 0.1 || endfreq > 10) {
        fprintf(stderr, "realfft: -startwidth 2..50, -endwidth 50..500, -endfreq 0.1..10\n");
        return 1;
    }
    for (int fi = first; fi < argc; fi++) {
        char root[4096], outroot[4400], outname[4500];
        const char *suf = split(argv[fi], root, sizeof(root));
        int isign = (suf && !strcmp(suf, "fft")) ? 1 : -1;
        if (fwd)
            isign = -1;
        if (inv)
            isign = 1;
        const char *outsuf = isign == -1 ? "fft" : "dat";
        const char *tag = red ? "_red" : "";
        if (outdir) {
            const char *base = strrchr(root, '/');
            snprintf(outroot, sizeof(outroot), "%s/%s%s", outdir, base ? base + 1 : root, tag);
        } else {
            snprintf(outroot, sizeof(outroot), "%s%s", root, tag);
        }
        snprintf(outname, sizeof(outname), "%s.%s", outroot, outsuf);
        infodata idata;
        double T = 0.0;
        if (red) {
            if (isign == 1) {
                printf("\nYou cannot remove red noise from an inverse FFT!\n\n");
                return 1;
            }
            if (accel_readinf(&idata, root)) {
                printf("\nCannot read '%s.inf', which is required by '-rednoise'!\n\n", root);
                return 1;
            }
            T = idata.N * idata.dt;
            if (strlen(outroot) >= sizeof(idata.name)) {
                printf("\nOutput file name '%s' is too long for a '.inf' file!\n\n", outroot);
                return 1;
            }
            strcpy(idata.name, outroot);
        }
        long long n = 0;
        float *data = (float *) read_all(argv[fi], sizeof(float), &n);
        if (!data) {
            fprintf(stderr, "realfft: cannot read '%s'\n", argv[fi]);
            return 1;
        }
        if (!even_len(n)) {
            fprintf(stderr, "\nrealfft() arrays lengths must be evenly divisible by 2.\n\n");
            return 1;
        }
        printf("\nPerforming in-core %s FFT on data:\n", isign == -1 ? "forward" : "inverse");
        printf("   Reading.\n   Transforming.\n");
        float *out = (float *) malloc(sizeof(float) * (size_t) n);
        int rc = isign == -1
                     ? accel_realfft_forward(data, n, (accel_fcomplex *) out, device)
                     : accel_realfft_inverse((const accel_fcomplex *) data, n, out, device);
        if (rc) {
            fprintf(stderr, "realfft: %s\n", accel_last_error());
            return 1;
        }
        if (red) {
            printf("   Removing red noise.\n");
            if (accel_deredden_gpu((accel_fcomplex *) out, n / 2, startwidth, endwidth, endfreq,
                                   T, device) < 0) {
                printf("\nRed noise removal failed.\n\n");
                return 1;
            }
        }
        printf("   Writing.\n");
        if (write_all(outname, out, sizeof(float), n)) {
            fprintf(stderr, "realfft: cannot write '%s'\n", outname);
            return 1;
        }
        if (red && accel_writeinf(&idata)) {
            fprintf(stderr, "realfft: %s\n", accel_last_error());
            return 1;
        }
        free(out);
        free(data);
        if (del)
            remove(argv[fi]);
        printf("\nFinished.  Wrote '%s'.\n\n", outname);
    }
    return 0;
}

#elif defined(TOOL_REDNOISE)

int main(int argc, char **argv)
{
    int startwidth = 6, endwidth = 100, inplace = 0, device = 0;
    double endfreq = 6.0;
    const char *in = NULL;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (!strcmp(a, "-startwidth") && i + 1 < argc) startwidth = atoi(argv[++i]);
        else if (!strcmp(a, "-endwidth") && i + 1 < argc) endwidth = atoi(argv[++i]);
        else if (!strcmp(a, "-endfreq") && i + 1 < argc) endfreq = atof(argv[++i]);
        else if (!strcmp(a, "-inplace")) inplace = 1;
        else if (!strcmp(a, "-device") && i + 1 < argc) device = atoi(argv[++i]);
        else if (a[0] == '-') {
            fprintf(stderr, "rednoise: unknown option '%s'\n", a);
            return 1;
        } else
            in = a;
    }
    if (!in) {
        printf("\nUsage:  rednoise [-startwidth s] [-endwidth e] [-endfreq f] [-inplace] "
               "[-device d] file.fft\n\n");
        return 1;
    }
    if (startwidth < 2 || startwidth > 50 || endwidth < 50 || endwidth > 500 ||
        endfreq < 0.1 || endfreq > 10) {
        fprintf(stderr, "rednoise: -startwidth 2..50, -endwidth 50..500, -endfreq 0.1..10\n");
        return 1;
    }
    printf("\n\n     Rednoise Removal Routine (MI355X)\n\n");
    char root[4096], newroot[4200], outname[4300];
    const char *suf = split(in, root, sizeof(root));
    if (!suf || strcmp(suf, "fft")) {
        printf("\nInput file ('%s') must be a fourier transform ('.fft')!\n\n", in);
        return 1;
    }
    snprintf(newroot, sizeof(newroot), "%s_red", root);
    snprintf(outname, sizeof(outname), "%s.fft", newroot);
    infodata idata;
    if (accel_readinf(&idata, root)) {
        fprintf(stderr, "rednoise: %s\n", accel_last_error());
        return 1;
    }
    const double T = idata.N * idata.dt;
    if (strlen(newroot) >= sizeof(idata.name)) {
        printf("\nOutput file name '%s' is too long for a '.inf' file!\n\n", newroot);
        return 1;
    }
    strcpy(idata.name, newroot);
    long long numbins = 0;
    accel_fcomplex *fft = (accel_fcomplex *) read_all(in, sizeof(accel_fcomplex), &numbins);
    if (!fft) {
        fprintf(stderr, "rednoise: cannot read '%s'\n", in);
        return 1;
    }
    if (inplace)
        printf("Dereddening '%s' in place.\n", in);
    long nw = accel_deredden_gpu(fft, numbins, startwidth, endwidth, endfreq, T, device);
    if (nw < 0) {
        printf("\nRednoise removal failed.\n\n");
        return 1;
    }
    if (inplace) {
        if (write_all(in, fft, sizeof(accel_fcomplex), numbins) || rename(in, outname)) {
            perror("\nError renaming the dereddened file");
            return 1;
        }
    } else if (write_all(outname, fft, sizeof(accel_fcomplex), numbins)) {
        fprintf(stderr, "rednoise: cannot write '%s'\n", outname);
        return 1;
    }
    if (accel_writeinf(&idata)) {
        fprintf(stderr, "rednoise: %s\n", accel_last_error());
        return 1;
    }
    printf("\nDone.  Rednoise removed from %ld of %lld points.\n\n", nw, numbins);
    printf("   Dereddened fft file: %s\n", outname);
    printf("Corresponding inf file: %s.inf\n\n", newroot);
    free(fft);
    return 0;
}

#elif defined(TOOL_ZAPBIRDS)

int main(int argc, char **argv)
{
    int zap = 0, constamp = 0, inzap = 0, outzap = 0;
    double baryv = 0.0;
    const char *zapfile = NULL, *in = NULL;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (!strcmp(a, "-zap")) zap = 1;
        else if (!strcmp(a, "-zapfile") && i + 1 < argc) zapfile = argv[++i];
        else if (!strcmp(a, "-in") && i + 1 < argc) { inzap = 1; ++i; }
        else if (!strcmp(a, "-out") && i + 1 < argc) { outzap = 1; ++i; }
        else if (!strcmp(a, "-baryv") && i + 1 < argc) baryv = atof(argv[++i]);
        else if (!strcmp(a, "-constamp")) constamp = 1;
        else if (a[0] == '-') {
            fprintf(stderr, "zapbirds: unknown option '%s'\n", a);
            return 1;
        } else
            in = a;
    }
    printf("\n\n     Automatic Birdie Zapping (MI355X)\n\n");
    if (!zap && !inzap && !outzap) {
        printf("You must specify '-in' and '-out' if you are not\n");
        printf("automatically zapping a file (with '-zap').\n\n");
        return 0;
    }
    if (!zap) {
        fprintf(stderr, "zapbirds: the interactive mode is not built; use -zap -zapfile\n");
        return 1;
    }
    if (baryv < -0.1 || baryv > 0.1) {
        fprintf(stderr, "zapbirds: -baryv -0.1..0.1\n");
        return 1;
    }
    char root[4096];
    const char *suf = in ? split(in, root, sizeof(root)) : NULL;
    if (!suf || strcmp(suf, "fft")) {
        printf("\nInput file ('%s') must be a FFT file ('.fft')!\n\n", in ? in : "");
        return 0;
    }
    infodata idata;
    if (accel_readinf(&idata, root)) {
        fprintf(stderr, "zapbirds: %s\n", accel_last_error());
        return 1;
    }
    printf("Examining data from '%s'.\n\n", in);
    if (!zapfile) {
        printf("You must specify a 'zapfile' containing freqs\n");
        printf("and widths if you want to write to the FFT file.\n\n");
        return 0;
    }
    const double T = idata.dt * idata.N;
    int nb = accel_read_birdies(zapfile, T, baryv, NULL, NULL, 0);
    if (nb < 0) {
        fprintf(stderr, "zapbirds: cannot read zapfile '%s'\n", zapfile);
        return 1;
    }
    double *lo = (double *) malloc(sizeof(double) * (nb ? nb : 1));
    double *hi = (double *) malloc(sizeof(double) * (nb ? nb : 1));
    accel_read_birdies(zapfile, T, baryv, lo, hi, nb);
    long long numbins = 0;
    accel_fcomplex *fft = (accel_fcomplex *) read_all(in, sizeof(accel_fcomplex), &numbins);
    if (!fft) {
        fprintf(stderr, "zapbirds: cannot read '%s'\n", in);
        return 1;
    }
    if (accel_zapbirds_fft(fft, numbins, idata.N, lo, hi, nb, constamp) ||
        write_all(in, fft, sizeof(accel_fcomplex), numbins)) {
        fprintf(stderr, "zapbirds: zapping '%s' failed\n", in);
        return 1;
    }
    printf("Zapped %d birdies from '%s'.\n\n", nb, in);
    free(fft);
    free(lo);
    free(hi);
    return 0;
}

#else
#error "build with one of TOOL_REALFFT, TOOL_REDNOISE, TOOL_ZAPBIRDS"
#endif
