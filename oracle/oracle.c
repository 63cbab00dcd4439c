/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h for what it restates
 * and which reference files it follows).  Built by oracle/Makefile into
 * oracle/liboracle.so; never linked into the product library.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

#define ORC_DIRECT_MAX 97

static const long double K2PI_L =
    6.283185307179586476925286766559005768394338798750211641949889L;

/* exp(sign*2*pi*i*m/n) with exact reduction of m mod n and octant symmetry,
   following the reduction of kernel/trig.c:57-80 (angle folded into
   [0, pi/4] using integer arithmetic before cosl/sinl). */
static void orc_cexpl(long m, long n, int sign, long double *c_out,
                      long double *s_out)
{
    __int128 mm = (__int128)m % n, N4, M4, q;
    unsigned oct = 0;
    long double th, c, s, t;
    if (mm < 0) mm += n;
    N4 = 4 * (__int128)n;
    M4 = 4 * mm;
    q = n;
    if (M4 > N4 - M4) { M4 = N4 - M4; oct |= 4; }
    if (M4 - q > 0) { M4 -= q; oct |= 2; }
    if (M4 > q - M4) { M4 = q - M4; oct |= 1; }
    th = (K2PI_L * (long double)M4) / (long double)N4;
    c = cosl(th);
    s = sinl(th);
    if (oct & 1) { t = c; c = s; s = t; }
    if (oct & 2) { t = c; c = -s; s = t; }
    if (oct & 4) { s = -s; }
    if (sign < 0) s = -s;
    *c_out = c;
    *s_out = s;
}

void orc_cexp(long m, long n, int sign, double *out2)
{
    long double c, s;
    orc_cexpl(m, n, sign, &c, &s);
    out2[0] = (double)c;
    out2[1] = (double)s;
}

/* ---- two instantiations of the FFT core ---- */
#define OR long double
#define OS(x) x##_ld
#include "oracle_core.h"
#undef OR
#undef OS

#define OR double
#define OS(x) x##_d
#include "oracle_core.h"
#undef OR
#undef OS

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- guru drivers ---- */

static int dims_ok(int rank, const orc_iodim *dims, int hrank,
                   const orc_iodim *hdims)
{
    int i;
    if (rank < 0 || hrank < 0 || rank > 32 || hrank > 32) return 0;
    for (i = 0; i < rank; ++i)
        if (dims[i].n <= 0) return 0;
    for (i = 0; i < hrank; ++i)
        if (hdims[i].n < 0) return 0;
    return 1;
}

/* offset of linear index lin over dims (row-major), using stride field */
static long off_of(long lin, int rank, const orc_iodim *d, int use_os)
{
    long off = 0;
    int i;
    for (i = rank - 1; i >= 0; --i) {
        long idx = lin % d[i].n;
        lin /= d[i].n;
        off += idx * (use_os ? d[i].os : d[i].is);
    }
    return off;
}

static long prod_n(int rank, const orc_iodim *d)
{
    long p = 1;
    int i;
    for (i = 0; i < rank; ++i) p *= d[i].n;
    return p;
}

#define DRIVERS(OR_, SFX)                                                     \
static void dft_##SFX(int rank, const orc_iodim *dims, int hrank,            \
                      const orc_iodim *hdims, const double *in, double *out,  \
                      int sign)                                               \
{                                                                             \
    long total = prod_n(rank, dims), H = prod_n(hrank, hdims), i;           \
    long n[33] = {0};                                                         \
    int d;                                                                    \
    cpx_##SFX *buf;                                                           \
    if (H == 0) return;                                                       \
    for (d = 0; d < rank; ++d) n[d] = dims[d].n;                              \
    buf = (cpx_##SFX *)malloc(sizeof(cpx_##SFX) * H * total);               \
    _Pragma("omp parallel for schedule(static)")                              \
    for (i = 0; i < H * total; ++i) {                                         \
        long h = i / total, e = i % total;                                    \
        long o = off_of(h, hrank, hdims, 0) + off_of(e, rank, dims, 0);       \
        buf[i].re = (OR_)in[2 * o];                                           \
        buf[i].im = (OR_)in[2 * o + 1];                                       \
    }                                                                         \
    separable_##SFX(buf, H, rank, n, 0, rank, sign);                          \
    _Pragma("omp parallel for schedule(static)")                              \
    for (i = 0; i < H * total; ++i) {                                         \
        long h = i / total, e = i % total;                                    \
        long o = off_of(h, hrank, hdims, 1) + off_of(e, rank, dims, 1);       \
        out[2 * o] = (double)buf[i].re;                                       \
        out[2 * o + 1] = (double)buf[i].im;                                   \
    }                                                                         \
    free(buf);                                                                \
}                                                                             \
                                                                              \
static void r2c_##SFX(int rank, const orc_iodim *dims, int hrank,            \
                      const orc_iodim *hdims, const double *in, double *out)  \
{                                                                             \
    long total = prod_n(rank, dims), H = prod_n(hrank, hdims), i;           \
    long n[33], nl, nh, tot_h;                                                \
    int d;                                                                    \
    cpx_##SFX *buf;                                                           \
    orc_iodim od[33];                                                         \
    if (H == 0) return;                                                       \
    for (d = 0; d < rank; ++d) { n[d] = dims[d].n; od[d] = dims[d]; }         \
    nl = rank ? n[rank - 1] : 1;                                              \
    nh = rank ? nl / 2 + 1 : 1;                                               \
    if (rank) od[rank - 1].n = nh;                                            \
    tot_h = rank ? total / nl * nh : 1;                                       \
    buf = (cpx_##SFX *)malloc(sizeof(cpx_##SFX) * H * total);               \
    _Pragma("omp parallel for schedule(static)")                              \
    for (i = 0; i < H * total; ++i) {                                         \
        long h = i / total, e = i % total;                                    \
        long o = off_of(h, hrank, hdims, 0) + off_of(e, rank, dims, 0);       \
        buf[i].re = (OR_)in[o];                                               \
        buf[i].im = 0;                                                        \
    }                                                                         \
    separable_##SFX(buf, H, rank, n, 0, rank, -1);                            \
    _Pragma("omp parallel for schedule(static)")                              \
    for (i = 0; i < H * tot_h; ++i) {                                         \
        long h = i / tot_h, e = i % tot_h;                                    \
        long row = e / nh, k = e % nh;                                        \
        long o = off_of(h, hrank, hdims, 1) + off_of(e, rank, od, 1);         \
        cpx_##SFX v = buf[h * total + row * nl + k];                          \
        out[2 * o] = (double)v.re;                                            \
        out[2 * o + 1] = (double)v.im;                                        \
    }                                                                         \
    free(buf);                                                                \
}                                                                             \
                                                                              \
static void c2r_##SFX(int rank, const orc_iodim *dims, int hrank,            \
                      const orc_iodim *hdims, const double *in, double *out)  \
{                                                                             \
    long H = prod_n(hrank, hdims), i, rows, nl, nh, tot_h;                    \
    long n[33] = {0};                                                         \
    int d;                                                                    \
    cpx_##SFX *buf;                                                           \
    orc_iodim id[33];                                                         \
    plan_##SFX *P;                                                            \
    if (H == 0) return;                                                       \
    if (rank == 0) {                                                          \
        for (i = 0; i < H; ++i)                                               \
            out[off_of(i, hrank, hdims, 1)] =                                 \
                in[2 * off_of(i, hrank, hdims, 0)];                           \
        return;                                                               \
    }                                                                         \
    for (d = 0; d < rank; ++d) { n[d] = dims[d].n; id[d] = dims[d]; }         \
    nl = n[rank - 1];                                                         \
    nh = nl / 2 + 1;                                                          \
    id[rank - 1].n = nh;                                                      \
    n[rank - 1] = nh;                                                         \
    tot_h = prod_n(rank, id);                                                 \
    rows = tot_h / nh;                                                        \
    buf = (cpx_##SFX *)malloc(sizeof(cpx_##SFX) * H * tot_h);               \
    _Pragma("omp parallel for schedule(static)")                              \
    for (i = 0; i < H * tot_h; ++i) {                                         \
        long h = i / tot_h, e = i % tot_h;                                    \
        long o = off_of(h, hrank, hdims, 0) + off_of(e, rank, id, 0);         \
        buf[i].re = (OR_)in[2 * o];                                           \
        buf[i].im = (OR_)in[2 * o + 1];                                       \
    }                                                                         \
    /* backward complex DFT over the leading dims (rank-geq2-rdft2 order) */  \
    separable_##SFX(buf, H, rank, n, 0, rank - 1, +1);                        \
    prepare_##SFX(nl);                                                        \
    P = get_plan_##SFX(nl);                                                   \
    _Pragma("omp parallel")                                                   \
    {                                                                         \
        cpx_##SFX *z = (cpx_##SFX *)malloc(sizeof(cpx_##SFX) * nl);          \
        cpx_##SFX *y = (cpx_##SFX *)malloc(sizeof(cpx_##SFX) * nl);          \
        long r, k;                                                            \
        _Pragma("omp for schedule(static)")                                   \
        for (r = 0; r < H * rows; ++r) {                                      \
            const cpx_##SFX *X = buf + r * nh;                                \
            long h = r / rows, rr = r % rows, j;                              \
            z[0].re = X[0].re; z[0].im = 0;                                   \
            for (k = 1; k < nh; ++k) {                                        \
                z[k] = X[k];                                                  \
                if (nl - k != k) {                                            \
                    z[nl - k].re = X[k].re;                                   \
                    z[nl - k].im = -X[k].im;                                  \
                } else {                                                      \
                    z[k].im = 0;                                              \
                }                                                             \
            }                                                                 \
            fft_##SFX(P, +1, z, 1, y);                                        \
            for (j = 0; j < nl; ++j) {                                        \
                long e = rr * nl + j;                                         \
                long o = off_of(h, hrank, hdims, 1) + off_of(e, rank, dims, 1); \
                out[o] = (double)y[j].re;                                     \
            }                                                                 \
        }                                                                     \
        free(z);                                                              \
        free(y);                                                              \
    }                                                                         \
    free(buf);                                                                \
}

/* Real-to-real kinds (doc/reference.texi:913-947 for the kinds, :2021-2350
   for their definitions).  Each 1-D transform is a complex DFT of the
   logical symmetric sequence the definition samples: the input is placed
   at its (half-)sample positions in a zero sequence of length N (2(n-1),
   2(n+1), 4n or 8n), mirrored with the kind's even / odd symmetry, and the
   outputs are read from the DFT at the kind's (half-)frequencies.  For
   example REDFT10 (Y_k = 2 sum x_j cos(pi (j+1/2) k / n)) puts x_j at 2j+1
   and 4n-2j-1 of a length-4n sequence, whose DFT at k is Y_k.  R2HC, HC2R
   and DHT are the n-point DFT read as halfcomplex, Hermitian-extended and
   Re - Im.  Multi-D: the separable product, one dim after another. */
static long r2r_logical(int kind, long n)
{
    switch (kind) {
    case 3: return 2 * (n - 1);          /* REDFT00 */
    case 7: return 2 * (n + 1);          /* RODFT00 */
    case 4: case 5: case 8: case 9: return 4 * n;
    case 6: case 10: return 8 * n;       /* REDFT11 / RODFT11 */
    default: return n;                   /* R2HC, HC2R, DHT */
    }
}

#define R2R_DRIVER(OR_, SFX)                                                  \
static void r2r_line_##SFX(int kind, long n, const OR_ *x, OR_ *y,           \
                           cpx_##SFX *z, cpx_##SFX *f, plan_##SFX *P)         \
{                                                                             \
    long N = r2r_logical(kind, n), j, k;                                      \
    int sign = kind == 1 ? +1 : -1;                                           \
    for (j = 0; j < N; ++j) { z[j].re = 0; z[j].im = 0; }                     \
    switch (kind) {                                                           \
    case 0: case 2:                                                           \
        for (j = 0; j < n; ++j) z[j].re = x[j];                               \
        break;                                                                \
    case 1:                                                                   \
        z[0].re = x[0];                                                       \
        for (k = 1; k < n; ++k) {                                             \
            if (2 * k < n) { z[k].re = x[k]; z[k].im = x[n - k]; }            \
            else if (2 * k == n) { z[k].re = x[k]; }                          \
            else { z[k].re = x[n - k]; z[k].im = -x[k]; }                     \
        }                                                                     \
        break;                                                                \
    case 3:                                                                   \
        for (j = 0; j < n; ++j) { z[j].re += x[j]; if (j && j < n - 1) z[N - j].re += x[j]; } \
        break;                                                                \
    case 7:                                                                   \
        for (j = 0; j < n; ++j) { z[j + 1].re = x[j]; z[N - j - 1].re = -x[j]; } \
        break;                                                                \
    case 5: case 6: case 9: case 10: {                                        \
        OR_ s = (kind == 9 || kind == 10) ? -1 : 1;                           \
        for (j = 0; j < n; ++j) { z[2 * j + 1].re = x[j]; z[N - 2 * j - 1].re = s * x[j]; } \
        break;                                                                \
    }                                                                         \
    case 4:                                                                   \
        z[0].re = x[0];                                                       \
        for (j = 1; j < n; ++j) { z[j].re = x[j]; z[N - j].re = x[j]; }       \
        break;                                                                \
    case 8:                                                                   \
        for (j = 0; j < n; ++j) {                                             \
            OR_ v = j == n - 1 ? x[j] / 2 : x[j];                             \
            z[j + 1].re += v;                                                 \
            z[N - j - 1].re -= v;                                             \
        }                                                                     \
        break;                                                                \
    }                                                                         \
    fft_##SFX(P, sign, z, 1, f);                                              \
    for (k = 0; k < n; ++k) {                                                 \
        switch (kind) {                                                       \
        case 0: y[k] = 2 * k <= n ? f[k].re : f[n - k].im; break;             \
        case 1: y[k] = f[k].re; break;                                        \
        case 2: y[k] = f[k].re - f[k].im; break;                              \
        case 3: case 5: y[k] = f[k].re; break;                                \
        case 4: case 6: y[k] = f[2 * k + 1].re; break;                        \
        case 7: case 9: y[k] = -f[k + 1].im; break;                           \
        case 8: case 10: y[k] = -f[2 * k + 1].im; break;                      \
        }                                                                     \
    }                                                                         \
}                                                                             \
                                                                              \
static void r2r_##SFX(int rank, const orc_iodim *dims, int hrank,            \
                      const orc_iodim *hdims, const int *kinds,              \
                      const double *in, double *out)                          \
{                                                                             \
    long total = prod_n(rank, dims), H = prod_n(hrank, hdims), i;           \
    int d;                                                                    \
    OR_ *buf;                                                                 \
    if (H == 0) return;                                                       \
    buf = (OR_ *)malloc(sizeof(OR_) * H * total);                            \
    for (i = 0; i < H * total; ++i) {                                         \
        long h = i / total, e = i % total;                                    \
        buf[i] = (OR_)in[off_of(h, hrank, hdims, 0) + off_of(e, rank, dims, 0)]; \
    }                                                                         \
    for (d = rank - 1; d >= 0; --d) {                                         \
        long nd = dims[d].n, inner = 1, lines, L, N = r2r_logical(kinds[d], nd); \
        int e;                                                                \
        plan_##SFX *P;                                                        \
        prepare_##SFX(N);                                                     \
        P = get_plan_##SFX(N);                                                \
        for (e = d + 1; e < rank; ++e) inner *= dims[e].n;                    \
        lines = H * (total / nd);                                             \
        _Pragma("omp parallel")                                               \
        {                                                                     \
            OR_ *tx = (OR_ *)malloc(sizeof(OR_) * nd);                        \
            OR_ *ty = (OR_ *)malloc(sizeof(OR_) * nd);                        \
            cpx_##SFX *z = (cpx_##SFX *)malloc(sizeof(cpx_##SFX) * N);        \
            cpx_##SFX *f = (cpx_##SFX *)malloc(sizeof(cpx_##SFX) * N);        \
            _Pragma("omp for schedule(static)")                               \
            for (L = 0; L < lines; ++L) {                                     \
                long h = L / (total / nd), r = L % (total / nd);              \
                long outer = r / inner, in_ = r % inner, j;                   \
                OR_ *base = buf + h * total + outer * nd * inner + in_;       \
                for (j = 0; j < nd; ++j) tx[j] = base[j * inner];             \
                r2r_line_##SFX(kinds[d], nd, tx, ty, z, f, P);                \
                for (j = 0; j < nd; ++j) base[j * inner] = ty[j];             \
            }                                                                 \
            free(tx); free(ty); free(z); free(f);                             \
        }                                                                     \
    }                                                                         \
    for (i = 0; i < H * total; ++i) {                                         \
        long h = i / total, e = i % total;                                    \
        out[off_of(h, hrank, hdims, 1) + off_of(e, rank, dims, 1)] = (double)buf[i]; \
    }                                                                         \
    free(buf);                                                                \
}

DRIVERS(long double, ld)
DRIVERS(double, d)
R2R_DRIVER(long double, ld)
R2R_DRIVER(double, d)

int orc_r2r(int prec, int rank, const orc_iodim *dims, int hrank,
            const orc_iodim *hdims, const int *kinds, const double *in,
            double *out)
{
    int i;
    if (!dims_ok(rank, dims, hrank, hdims)) return -1;
    for (i = 0; i < rank; ++i) {
        if (kinds[i] < 0 || kinds[i] > 10) return -1;
        if (kinds[i] == 3 && dims[i].n == 1) return -1; /* REDFT00 of size 1 */
    }
    if (prec == ORC_LD)
        r2r_ld(rank, dims, hrank, hdims, kinds, in, out);
    else
        r2r_d(rank, dims, hrank, hdims, kinds, in, out);
    return 0;
}

int orc_dft(int prec, int rank, const orc_iodim *dims, int hrank,
            const orc_iodim *hdims, const double *in, double *out, int sign)
{
    if (!dims_ok(rank, dims, hrank, hdims) || (sign != -1 && sign != 1))
        return -1;
    if (prec == ORC_LD)
        dft_ld(rank, dims, hrank, hdims, in, out, sign);
    else
        dft_d(rank, dims, hrank, hdims, in, out, sign);
    return 0;
}

int orc_r2c(int prec, int rank, const orc_iodim *dims, int hrank,
            const orc_iodim *hdims, const double *in, double *out)
{
    if (!dims_ok(rank, dims, hrank, hdims)) return -1;
    if (prec == ORC_LD)
        r2c_ld(rank, dims, hrank, hdims, in, out);
    else
        r2c_d(rank, dims, hrank, hdims, in, out);
    return 0;
}

int orc_c2r(int prec, int rank, const orc_iodim *dims, int hrank,
            const orc_iodim *hdims, const double *in, double *out)
{
    if (!dims_ok(rank, dims, hrank, hdims)) return -1;
    if (prec == ORC_LD)
        c2r_ld(rank, dims, hrank, hdims, in, out);
    else
        c2r_d(rank, dims, hrank, hdims, in, out);
    return 0;
}

/* ---- independent brute-force checkers (no code shared with the FFT) ---- */

typedef struct { long n, b; long double *lo, *hi; } tbl2;

static void tbl2_make(tbl2 *t, long n, int sign)
{
    long e;
    t->n = n;
    t->b = 1;
    while (t->b * t->b < n) t->b++;
    t->lo = (long double *)malloc(sizeof(long double) * 2 * t->b);
    t->hi = (long double *)malloc(sizeof(long double) * 2 * (n / t->b + 1));
    for (e = 0; e < t->b; ++e)
        orc_cexpl(e, n, sign, &t->lo[2 * e], &t->lo[2 * e + 1]);
    for (e = 0; e <= n / t->b; ++e)
        orc_cexpl(e * t->b, n, sign, &t->hi[2 * e], &t->hi[2 * e + 1]);
}

static void tbl2_free(tbl2 *t)
{
    free(t->lo);
    free(t->hi);
}

static inline void tbl2_get(const tbl2 *t, long e, long double *re,
                            long double *im)
{
    const long double *a = t->lo + 2 * (e % t->b), *b = t->hi + 2 * (e / t->b);
    *re = a[0] * b[0] - a[1] * b[1];
    *im = a[0] * b[1] + a[1] * b[0];
}

static void bin_sum(const tbl2 *t, const double *in, long k, double *out)
{
    long n = t->n, j, e = 0;
    long double re = 0, im = 0;
    k %= n;
    if (k < 0) k += n;
    for (j = 0; j < n; ++j) {
        long double xr = in[2 * j], xi = in[2 * j + 1], wr, wi;
        tbl2_get(t, e, &wr, &wi);
        re += xr * wr - xi * wi;
        im += xr * wi + xi * wr;
        e += k;
        if (e >= n) e -= n;
    }
    out[0] = (double)re;
    out[1] = (double)im;
}

void orc_dft_direct(long n, int sign, const double *in, double *out)
{
    tbl2 t;
    long k;
    tbl2_make(&t, n, sign);
#pragma omp parallel for schedule(dynamic, 16)
    for (k = 0; k < n; ++k) bin_sum(&t, in, k, out + 2 * k);
    tbl2_free(&t);
}

void orc_dft_bins(long n, int sign, const double *in, const long *bins,
                  long nbins, double *out)
{
    tbl2 t;
    long i;
    tbl2_make(&t, n, sign);
#pragma omp parallel for schedule(dynamic, 1)
    for (i = 0; i < nbins; ++i) bin_sum(&t, in, bins[i], out + 2 * i);
    tbl2_free(&t);
}

/* Sampled bins of a rank-r DFT of a row-major array (n[0] slowest): bin
   (k_0, ..., k_{r-1}) = sum_j x_j prod_d w_{n_d}^{j_d k_d}, summed one
   dimension at a time from the fastest (separability), in long double with
   exactly reduced exponents.  `real` says the input holds reals (imaginary
   parts zero); bins holds nbins rank-tuples. */
void orc_dft_bins_nd(int rank, const long *n, int sign, const double *in, int real,
                     const long *bins, long nbins, double *out)
{
    long b;
    tbl2 t[8];
    int d;
    if (rank < 1 || rank > 8) return;
    for (d = 0; d < rank; ++d) tbl2_make(&t[d], n[d], sign);
    for (b = 0; b < nbins; ++b) {
        const long *k = bins + (long)rank * b;
        long outer = 1, m, i;
        long double *acc, *nxt;
        for (d = 0; d < rank - 1; ++d) outer *= n[d];
        acc = (long double *)malloc(sizeof(long double) * 2 * outer);
        /* fastest dimension, straight from the input */
#pragma omp parallel for schedule(static)
        for (i = 0; i < outer; ++i) {
            const long nl = n[rank - 1];
            long kk = k[rank - 1] % nl, e = 0, j;
            long double re = 0, im = 0;
            if (kk < 0) kk += nl;
            for (j = 0; j < nl; ++j) {
                long double xr, xi, wr, wi;
                if (real) {
                    xr = in[i * nl + j];
                    xi = 0;
                } else {
                    xr = in[2 * (i * nl + j)];
                    xi = in[2 * (i * nl + j) + 1];
                }
                tbl2_get(&t[rank - 1], e, &wr, &wi);
                re += xr * wr - xi * wi;
                im += xr * wi + xi * wr;
                e += kk;
                if (e >= nl) e -= nl;
            }
            acc[2 * i] = re;
            acc[2 * i + 1] = im;
        }
        /* the remaining dimensions on the partial sums */
        for (d = rank - 2; d >= 0; --d) {
            const long nd = n[d];
            long kk = k[d] % nd;
            if (kk < 0) kk += nd;
            m = outer / nd;
            nxt = (long double *)malloc(sizeof(long double) * 2 * (m ? m : 1));
            for (i = 0; i < m; ++i) {
                long e = 0, j;
                long double re = 0, im = 0;
                for (j = 0; j < nd; ++j) {
                    long double xr = acc[2 * (i * nd + j)], xi = acc[2 * (i * nd + j) + 1], wr, wi;
                    tbl2_get(&t[d], e, &wr, &wi);
                    re += xr * wr - xi * wi;
                    im += xr * wi + xi * wr;
                    e += kk;
                    if (e >= nd) e -= nd;
                }
                nxt[2 * i] = re;
                nxt[2 * i + 1] = im;
            }
            free(acc);
            acc = nxt;
            outer = m;
        }
        out[2 * b] = (double)acc[0];
        out[2 * b + 1] = (double)acc[1];
        free(acc);
    }
    for (d = 0; d < rank; ++d) tbl2_free(&t[d]);
}

/* ---- counter-based inputs ---- */

uint64_t orc_splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double u01(uint64_t seed, uint64_t stream, uint64_t i)
{
    uint64_t z = orc_splitmix64(seed ^ (stream << 40) ^ i);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

void orc_fill_uniform(double *x, long count, uint64_t seed, uint64_t stream)
{
    long i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < count; ++i)
        x[i] = 2.0 * u01(seed, stream, (uint64_t)i) - 1.0;
}

void orc_fill_gauss(double *x, long count, uint64_t seed, uint64_t stream)
{
    long i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < count; i += 2) {
        double u1 = u01(seed, stream, (uint64_t)i);
        double u2 = u01(seed, stream, (uint64_t)i + 1);
        double r = sqrt(-2.0 * log(1.0 - u1));
        x[i] = r * cos(6.283185307179586 * u2);
        if (i + 1 < count) x[i + 1] = r * sin(6.283185307179586 * u2);
    }
}
