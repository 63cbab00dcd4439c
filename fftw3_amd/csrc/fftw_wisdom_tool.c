/*
 * fftw-wisdom for the MI355X engine: plans a list of transforms on the
 * device (FFTW_PATIENT unless told otherwise) and writes the resulting
 * wisdom, so that later programs plan with FFTW_WISDOM_ONLY or pick up the
 * tuned decisions immediately.
 *
 * Command line and problem syntax follow the reference tool
 * (tools/fftw-wisdom.c:72-130 options and help, :134-147 canonical sizes,
 * :291-330 problem list and time limit) and libbench2's problem strings
 * (libbench2/problem.c:229-318):
 *   <problem>  = <prefix>* <n1>[<kind>][x<n2>[<kind>]...] [*<howmany>]
 *   <prefix>   = c (complex) | r (real: r2c forward, c2r backward)
 *              | k (real-to-real) | i (in place) | o (out of place)
 *              | f / - (forward) | b / + (backward)
 *              | / (split real and imaginary arrays)
 *   <kind>     = r2r kind of the preceding dim (libbench2/problem.c:152-195):
 *                f R2HC (default), b HC2R, h DHT, e00/e01/e10/e11 REDFTab,
 *                o00/o01/o10/o11 RODFTab
 * The engine keeps both precisions in one wisdom store; -s (or running the
 * program as fftwf-wisdom) plans with the fftwf_ planners.
 *
 * This tool talks to the engine only through the drop-in C ABI
 * (include/fftw3.h) and the device-memory helpers (include/fftw3_amd.h).
 */
#include <ctype.h>
#include <getopt.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "../../include/fftw3.h"
#include "../../include/fftw3_amd.h"

#define MAXR 8

typedef struct {
    char text[64];
    int real, inplace, sign, split, r2r;
    int rank, n[MAXR];
    int kind[MAXR];            /* r2r kinds (fftw_r2r_kind) */
    int howmany;
    double size;               /* points, for ordering */
} problem;

/* powers of two and ten up to 2^20 / 10^6, and the square / cubic sizes
   of the reference's canonical set (tools/fftw-wisdom.c:134-147) */
static const char *canonical_sizes[] = {
    "1", "2", "4", "8", "16", "32", "64", "128", "256", "512", "1024", "2048", "4096", "8192",
    "16384", "32768", "65536", "131072", "262144", "524288", "1048576",
    "10", "100", "1000", "10000", "100000", "1000000",
    "2x2", "4x4", "8x8", "10x10", "16x16", "32x32", "64x64", "100x100", "128x128", "256x256",
    "512x512", "1000x1000", "1024x1024",
    "2x2x2", "4x4x4", "8x8x8", "10x10x10", "16x16x16", "32x32x32", "64x64x64", "100x100x100"};
static const char *canonical_types[] = {"cof", "cob", "cif", "cib", "rof", "rob", "rif", "rib"};

static int parse_problem(const char *s, problem *p)
{
    const char *q = s;
    char *end;
    long v;
    memset(p, 0, sizeof(*p));
    snprintf(p->text, sizeof(p->text), "%s", s);
    p->sign = -1;
    p->howmany = 1;
    for (;; ++q) {
        int ch = tolower((unsigned char)*q);
        if (ch == 'c') p->real = p->r2r = 0;
        else if (ch == 'r') { p->real = 1; p->r2r = 0; }
        else if (ch == 'k') { p->r2r = 1; p->real = 0; }
        else if (ch == 'i') p->inplace = 1;
        else if (ch == 'o') p->inplace = 0;
        else if (ch == 'f' || ch == '-') p->sign = -1;
        else if (ch == 'b' || ch == '+') p->sign = 1;
        else if (ch == '/') p->split = 1;
        else break;
    }
    for (;;) {
        if (p->rank == MAXR) return -1;
        if (!isdigit((unsigned char)*q)) return -1;
        v = strtol(q, &end, 10);
        if (v <= 0 || v > (1L << 30)) return -1;
        p->n[p->rank] = (int)v;
        p->kind[p->rank++] = FFTW_R2HC;
        q = end;
        if (p->r2r) {
            int ch = tolower((unsigned char)*q);
            if (ch == 'f') { ++q; }
            else if (ch == 'b') { p->kind[p->rank - 1] = FFTW_HC2R; ++q; }
            else if (ch == 'h') { p->kind[p->rank - 1] = FFTW_DHT; ++q; }
            else if (ch == 'e' || ch == 'o') {
                static const char *ab[] = {"00", "01", "10", "11"};
                int t, hit = -1;
                for (t = 0; t < 4; ++t)
                    if (!strncmp(q + 1, ab[t], 2) && !isdigit((unsigned char)q[3])) hit = t;
                if (hit < 0) return -1;
                p->kind[p->rank - 1] = (ch == 'e' ? FFTW_REDFT00 : FFTW_RODFT00) + hit;
                q += 3;
            }
        }
        if (*q == 'x' || *q == 'X') {
            ++q;
            continue;
        }
        break;
    }
    if (*q == '*') {
        ++q;
        if (!isdigit((unsigned char)*q)) return -1;
        v = strtol(q, &end, 10);
        if (v <= 0 || v > (1L << 30)) return -1;
        p->howmany = (int)v;
        q = end;
    }
    if (*q) return -1; /* trailing characters: strides, ... */
    if (p->split && (p->real || p->r2r)) return -1;
    p->size = p->howmany;
    for (v = 0; v < p->rank; ++v) p->size *= p->n[v];
    return 0;
}

static int by_size(const void *a, const void *b)
{
    double x = ((const problem *)a)->size, y = ((const problem *)b)->size;
    return x < y ? -1 : x > y;
}

/* Plans one problem on freshly allocated device arrays; 0 on success. */
static int plan_problem(const problem *p, int single, unsigned flags)
{
    size_t rs = single ? sizeof(float) : sizeof(double), cs = 2 * rs;
    long long N = 1, Nh, padded;
    int i, ok = 0, dist_r, dist_c;
    void *in = 0, *out = 0, *in2 = 0, *out2 = 0, *plan = 0;
    for (i = 0; i < p->rank; ++i) N *= p->n[i];
    Nh = N / p->n[p->rank - 1] * (p->n[p->rank - 1] / 2 + 1);
    padded = 2 * Nh;
    if (p->r2r) {
        size_t bytes = (size_t)N * p->howmany * rs;
        in = fftw_malloc(bytes);
        out = p->inplace ? in : fftw_malloc(bytes);
    } else if (!p->real) {
        size_t bytes = (size_t)N * p->howmany * (p->split ? rs : cs);
        in = fftw_malloc(bytes);
        out = p->inplace ? in : fftw_malloc(bytes);
        if (p->split) {
            in2 = fftw_malloc(bytes);
            out2 = p->inplace ? in2 : fftw_malloc(bytes);
        }
    } else {
        size_t rbytes = (size_t)(p->inplace ? padded : N) * p->howmany * rs;
        size_t cbytes = (size_t)Nh * p->howmany * cs;
        in = fftw_malloc(p->sign < 0 ? rbytes : cbytes);
        out = p->inplace ? in : fftw_malloc(p->sign < 0 ? cbytes : rbytes);
    }
    if (!in || !out || (p->split && (!in2 || !out2))) goto done;
    dist_r = (int)(p->inplace ? padded : N);
    dist_c = (int)Nh;
    if (p->r2r) {
        if (single)
            plan = fftwf_plan_many_r2r(p->rank, p->n, p->howmany, in, 0, 1, (int)N, out, 0, 1,
                                       (int)N, (const fftwf_r2r_kind *)p->kind, flags);
        else
            plan = fftw_plan_many_r2r(p->rank, p->n, p->howmany, in, 0, 1, (int)N, out, 0, 1,
                                      (int)N, (const fftw_r2r_kind *)p->kind, flags);
    } else if (p->split) {
        fftw_iodim dims[MAXR], hd;
        long long s = 1;
        for (i = p->rank - 1; i >= 0; --i) {
            dims[i].n = p->n[i];
            dims[i].is = dims[i].os = (int)s;
            s *= p->n[i];
        }
        hd.n = p->howmany;
        hd.is = hd.os = (int)N;
        /* forward on (re, im); backward by swapping the pointers */
        if (single)
            plan = p->sign < 0
                       ? (void *)fftwf_plan_guru_split_dft(p->rank, (fftwf_iodim *)dims, 1,
                                                           (fftwf_iodim *)&hd, in, in2, out, out2,
                                                           flags)
                       : (void *)fftwf_plan_guru_split_dft(p->rank, (fftwf_iodim *)dims, 1,
                                                           (fftwf_iodim *)&hd, in2, in, out2, out,
                                                           flags);
        else
            plan = p->sign < 0 ? (void *)fftw_plan_guru_split_dft(p->rank, dims, 1, &hd, in, in2,
                                                                  out, out2, flags)
                               : (void *)fftw_plan_guru_split_dft(p->rank, dims, 1, &hd, in2, in,
                                                                  out2, out, flags);
    } else if (!p->real) {
        if (single)
            plan = fftwf_plan_many_dft(p->rank, p->n, p->howmany, in, 0, 1, (int)N, out, 0, 1,
                                       (int)N, p->sign, flags);
        else
            plan = fftw_plan_many_dft(p->rank, p->n, p->howmany, in, 0, 1, (int)N, out, 0, 1,
                                      (int)N, p->sign, flags);
    } else if (p->sign < 0) {
        if (single)
            plan = fftwf_plan_many_dft_r2c(p->rank, p->n, p->howmany, in, 0, 1, dist_r, out, 0, 1,
                                           dist_c, flags);
        else
            plan = fftw_plan_many_dft_r2c(p->rank, p->n, p->howmany, in, 0, 1, dist_r, out, 0, 1,
                                          dist_c, flags);
    } else {
        if (single)
            plan = fftwf_plan_many_dft_c2r(p->rank, p->n, p->howmany, in, 0, 1, dist_c, out, 0, 1,
                                           dist_r, flags);
        else
            plan = fftw_plan_many_dft_c2r(p->rank, p->n, p->howmany, in, 0, 1, dist_c, out, 0, 1,
                                          dist_r, flags);
    }
    if (plan) {
        ok = 1;
        if (single) fftwf_destroy_plan(plan);
        else fftw_destroy_plan(plan);
    }
done:
    if (out2 && out2 != in2) fftw_free(out2);
    if (in2) fftw_free(in2);
    if (out && out != in) fftw_free(out);
    if (in) fftw_free(in);
    return ok ? 0 : -1;
}

static void help(FILE *f, const char *prog)
{
    fprintf(f,
            "Usage: %s [options] [sizes]\n"
            "    Create wisdom (pre-planned, device-tuned transforms) for the given sizes\n"
            "    on the MI355X engine, writing wisdom to stdout (or to a file, using -o).\n"
            "\nOptions:\n"
            "                   -h, --help: print this help\n"
            "                -V, --version: print version information\n"
            "                -v, --verbose: verbose output\n"
            "              -c, --canonical: plan the canonical set of sizes\n"
            "     -t <h>, --time-limit=<h>: time limit in hours (default: 0, no limit)\n"
            "  -o FILE, --output-file=FILE: output to FILE instead of stdout\n"
            "                -m, --measure: plan in MEASURE mode (PATIENT is default)\n"
            "               -e, --estimate: plan in ESTIMATE mode (not recommended)\n"
            "             -x, --exhaustive: plan in EXHAUSTIVE mode (may be slow)\n"
            "       -n, --no-system-wisdom: don't read the system wisdom file\n"
            "  -w FILE, --wisdom-file=FILE: read wisdom from FILE (stdin if -)\n"
            "                 -s, --single: plan single-precision (fftwf_) transforms\n"
            "                   -l, --list: print the parsed problem list and exit\n"
            "\nSize syntax: <type><inplace><direction><geometry>[*<howmany>]\n"
            "      <type> = c/r/k for complex/real (r2c forward, c2r backward)/r2r;\n"
            "               a '/' prefix asks for split real and imaginary arrays\n"
            "   <inplace> = i/o for in/out-of place\n"
            " <direction> = f/b for forward/backward, omitted for k transforms\n"
            "  <geometry> = <n1>[x<n2>[x...]], e.g. 10x12x14\n"
            "               -- for k transforms, after each dimension is a <kind>:\n"
            "                     <kind> = f/b/h/e00/e01/e10/e11/o00/o01/o10/o11\n"
            "                              for R2HC/HC2R/DHT/REDFT00/.../RODFT11\n",
            prog);
}

int main(int argc, char **argv)
{
    static const struct option opts[] = {
        {"help", no_argument, 0, 'h'},         {"version", no_argument, 0, 'V'},
        {"verbose", no_argument, 0, 'v'},      {"canonical", no_argument, 0, 'c'},
        {"time-limit", required_argument, 0, 't'}, {"output-file", required_argument, 0, 'o'},
        {"impatient", no_argument, 0, 'i'},    {"measure", no_argument, 0, 'm'},
        {"estimate", no_argument, 0, 'e'},     {"exhaustive", no_argument, 0, 'x'},
        {"no-system-wisdom", no_argument, 0, 'n'}, {"wisdom-file", required_argument, 0, 'w'},
        {"single", no_argument, 0, 's'},       {"list", no_argument, 0, 'l'},
        {0, 0, 0, 0}};
    const char *base = strrchr(argv[0], '/') ? strrchr(argv[0], '/') + 1 : argv[0];
    int verbose = 0, canonical = 0, impatient = 0, system_wisdom = 1, list = 0;
    int single = strncmp(base, "fftwf", 5) == 0;
    unsigned extra = 0;
    double hours = 0;
    const char *ofile = 0;
    problem *probs = 0;
    int np = 0, cap = 0, c, i, j, failed = 0;
    time_t begin;
    FILE *of;

    while ((c = getopt_long(argc, argv, "hVvct:o:imexnw:sl", opts, 0)) != -1) {
        switch (c) {
        case 'h': help(stdout, base); return 0;
        case 'V': printf("fftw-wisdom tool for %s\n", fftw_version); return 0;
        case 'v': verbose = 1; break;
        case 'c': canonical = 1; break;
        case 't': hours = atof(optarg); break;
        case 'o': ofile = strcmp(optarg, "-") ? optarg : 0; break;
        case 'i':
        case 'm': impatient = 1; break;
        case 'e': extra |= FFTW_ESTIMATE; break;
        case 'x': extra |= FFTW_EXHAUSTIVE; break;
        case 'n': system_wisdom = 0; break;
        case 's': single = 1; break;
        case 'l': list = 1; break;
        case 'w': {
            FILE *w = strcmp(optarg, "-") ? fopen(optarg, "r") : stdin;
            if (!w) {
                fprintf(stderr, "fftw-wisdom: error opening \"%s\"\n", optarg);
                return 1;
            }
            if (!fftw_import_wisdom_from_file(w)) {
                fprintf(stderr, "fftw-wisdom: error reading wisdom from \"%s\"\n", optarg);
                return 1;
            }
            if (w != stdin) fclose(w);
            break;
        }
        default: help(stderr, base); return 1;
        }
    }

#define ADD(str)                                                                          \
    do {                                                                                  \
        if (np == cap) {                                                                  \
            cap = cap ? 2 * cap : 64;                                                     \
            probs = (problem *)realloc(probs, sizeof(problem) * cap);                     \
        }                                                                                 \
        if (parse_problem((str), probs + np)) {                                           \
            fprintf(stderr, "fftw-wisdom: cannot parse problem \"%s\"\n", (str));         \
            return 1;                                                                     \
        }                                                                                 \
        ++np;                                                                             \
    } while (0)

    if (canonical)
        for (i = 0; i < (int)(sizeof(canonical_sizes) / sizeof(*canonical_sizes)); ++i)
            for (j = 0; j < (int)(sizeof(canonical_types) / sizeof(*canonical_types)); ++j) {
                char ps[64];
                /* multi-dimensional sizes only in place */
                if (strchr(canonical_sizes[i], 'x') && strchr(canonical_types[j], 'o')) continue;
                snprintf(ps, sizeof(ps), "%s%s", canonical_types[j], canonical_sizes[i]);
                ADD(ps);
            }
    for (; optind < argc; ++optind) {
        if (!strcmp(argv[optind], "-")) {
            char s[1025];
            while (fscanf(stdin, "%1024s", s) == 1) ADD(s);
        } else {
            ADD(argv[optind]);
        }
    }
    qsort(probs, (size_t)np, sizeof(problem), by_size);

    if (list) {
        for (i = 0; i < np; ++i) {
            static const char *kn[] = {"r2hc", "hc2r", "dht", "redft00", "redft01", "redft10",
                                       "redft11", "rodft00", "rodft01", "rodft10", "rodft11"};
            if (probs[i].r2r)
                printf("%s r2r %s rank=%d n=", probs[i].text,
                       probs[i].inplace ? "in-place" : "out-of-place", probs[i].rank);
            else
                printf("%s %s %s %s%s rank=%d n=", probs[i].text,
                       probs[i].real ? "real" : "complex",
                       probs[i].inplace ? "in-place" : "out-of-place",
                       probs[i].sign < 0 ? "forward" : "backward", probs[i].split ? " split" : "",
                       probs[i].rank);
            for (j = 0; j < probs[i].rank; ++j) printf("%s%d", j ? "x" : "", probs[i].n[j]);
            if (probs[i].r2r) {
                printf(" kind=");
                for (j = 0; j < probs[i].rank; ++j) printf("%s%s", j ? "," : "", kn[probs[i].kind[j]]);
            }
            printf(" howmany=%d\n", probs[i].howmany);
        }
        free(probs);
        return 0;
    }

    if (fftw_amd_init()) {
        fprintf(stderr, "fftw-wisdom: %s\n", fftw_amd_last_error());
        return 1;
    }
    if (system_wisdom && !fftw_import_system_wisdom() && verbose)
        fprintf(stderr, "fftw-wisdom: system-wisdom import failed\n");
    of = ofile ? fopen(ofile, "w") : stdout;
    if (!of) {
        fprintf(stderr, "fftw-wisdom: error creating \"%s\"\n", ofile);
        return 1;
    }
    begin = time(0);
    for (i = 0; i < np; ++i) {
        unsigned flags = extra ? extra : (impatient ? FFTW_MEASURE : FFTW_PATIENT);
        if (hours > 0 && (time(0) - begin) / 3600.0 >= hours) break;
        if (verbose) fprintf(stderr, "Planning transform: %s\n", probs[i].text);
        if (plan_problem(probs + i, single, flags)) {
            fprintf(stderr, "fftw-wisdom: could not plan %s: %s\n", probs[i].text,
                    fftw_amd_last_error());
            ++failed;
        }
    }
    fftw_export_wisdom_to_file(of);
    if (of != stdout) fclose(of);
    free(probs);
    return failed ? 2 : 0;
}
