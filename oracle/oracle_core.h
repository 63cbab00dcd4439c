/*
 * Oracle FFT core, instantiated twice by oracle.c:
 *   OR  = internal real type (long double or double)
 *   OS(x) = suffix macro for the instantiation's names
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Algorithm (textbook mixed-radix decimation in time, written from the DFT
 * definition, doc/reference.texi:1872-1902):
 *   n = p * m, p = smallest prime factor of n
 *   Y[k' + m*s] = sum_q W_p^{q*s} * (W_n^{q*k'} * FFT_m(x[q::p])[k'])
 * Each p-point DFT is a direct O(p^2) sum for p <= ORC_DIRECT_MAX, and a
 * Bluestein convolution (dft/bluestein.c:37-128 conventions: chirp
 * exp(i*pi*m^2/p), M-point power-of-two FFTs) for larger primes.
 * Twiddles W_n^e = exp(-2*pi*i*e/n) come from a two-level table of
 * long-double values (W_n^e = lo[e % b] * hi[e / b]); every entry is
 * computed by orc_cexpl() with exact integer octant reduction
 * (kernel/trig.c:57-80).
 */

typedef struct { OR re, im; } OS(cpx);

typedef struct OS(plan_s) {
    long n;
    int nf;
    long f[64];                 /* prime factors, smallest first */
    long b;                     /* two-level table split */
    OS(cpx) *lo, *hi;           /* W_n^e = lo[e % b] * hi[e / b] */
    /* Bluestein (only when n is a prime > ORC_DIRECT_MAX) */
    long M;
    struct OS(plan_s) *sub;
    OS(cpx) *chirp;             /* exp(-i*pi*m^2/n), m < n */
    OS(cpx) *bf, *bb;           /* FFT_M of the conjugate chirp / M, per sign */
    struct OS(plan_s) *next;
} OS(plan);

static OS(plan) *OS(plans) = 0;

static inline OS(cpx) OS(cmul)(OS(cpx) a, OS(cpx) b)
{
    OS(cpx) r;
    r.re = a.re * b.re - a.im * b.im;
    r.im = a.re * b.im + a.im * b.re;
    return r;
}

/* W_n^e * (sign>0 ? conj : id), e any integer */
static inline OS(cpx) OS(tw)(const OS(plan) *P, long e, long n_sub_mul, int sign)
{
    long n = P->n;
    long ee = (e >= 0 && e < (1L << 31) && n_sub_mul < (1L << 31))
                  ? (e * n_sub_mul) % n
                  : (long)(((__int128)e * n_sub_mul) % n);
    if (ee < 0) ee += n;
    OS(cpx) w = OS(cmul)(P->lo[ee % P->b], P->hi[ee / P->b]);
    if (sign > 0) w.im = -w.im;
    return w;
}

static void OS(fft)(const OS(plan) *P, int sign, const OS(cpx) *x, long xs,
                    OS(cpx) *y);

static OS(plan) *OS(get_plan)(long n);

static OS(plan) *OS(make_plan)(long n)
{
    OS(plan) *P = (OS(plan) *)calloc(1, sizeof(OS(plan)));
    long r = n, p, e;
    P->n = n;
    while (r > 1) {
        for (p = 2; p * p <= r; ++p)
            if (r % p == 0) break;
        if (p * p > r) p = r;
        P->f[P->nf++] = p;
        r /= p;
    }
    P->b = 1;
    while (P->b * P->b < n) P->b++;
    P->lo = (OS(cpx) *)malloc(sizeof(OS(cpx)) * P->b);
    P->hi = (OS(cpx) *)malloc(sizeof(OS(cpx)) * (n / P->b + 1));
    for (e = 0; e < P->b; ++e) {
        long double c, s;
        orc_cexpl(e, n, -1, &c, &s);
        P->lo[e].re = (OR)c; P->lo[e].im = (OR)s;
    }
    for (e = 0; e <= n / P->b; ++e) {
        long double c, s;
        orc_cexpl(e * P->b, n, -1, &c, &s);
        P->hi[e].re = (OR)c; P->hi[e].im = (OR)s;
    }
    if (P->nf == 1 && n > ORC_DIRECT_MAX) {
        long M = 1, m;
        OS(cpx) *tmp;
        while (M < 2 * n - 1) M <<= 1;
        P->M = M;
        P->sub = OS(get_plan)(M);
        P->chirp = (OS(cpx) *)malloc(sizeof(OS(cpx)) * n);
        for (m = 0; m < n; ++m) {
            long double c, s;
            long msq = (long)(((__int128)m * m) % (2 * n));
            orc_cexpl(msq, 2 * n, -1, &c, &s);       /* exp(-i*pi*m^2/n) */
            P->chirp[m].re = (OR)c; P->chirp[m].im = (OR)s;
        }
        P->bf = (OS(cpx) *)calloc(M, sizeof(OS(cpx)));
        P->bb = (OS(cpx) *)calloc(M, sizeof(OS(cpx)));
        tmp = (OS(cpx) *)calloc(M, sizeof(OS(cpx)));
        /* forward sign: c_m = chirp, b_m = conj(c_m) for |m| < n */
        for (m = 0; m < M; ++m) tmp[m].re = tmp[m].im = 0;
        for (m = 0; m < n; ++m) {
            OS(cpx) v = P->chirp[m];
            v.im = -v.im;
            tmp[m] = v;
            if (m) tmp[M - m] = v;
        }
        OS(fft)(P->sub, -1, tmp, 1, P->bf);
        /* backward sign: c_m = conj(chirp), b_m = chirp */
        for (m = 0; m < M; ++m) tmp[m].re = tmp[m].im = 0;
        for (m = 0; m < n; ++m) {
            tmp[m] = P->chirp[m];
            if (m) tmp[M - m] = P->chirp[m];
        }
        OS(fft)(P->sub, -1, tmp, 1, P->bb);
        for (m = 0; m < M; ++m) {
            P->bf[m].re /= (OR)M; P->bf[m].im /= (OR)M;
            P->bb[m].re /= (OR)M; P->bb[m].im /= (OR)M;
        }
        free(tmp);
    }
    return P;
}

/* Not thread-safe: callers create every plan they need before entering a
   parallel region (OS(prepare)). */
static OS(plan) *OS(get_plan)(long n)
{
    OS(plan) *P;
    for (P = OS(plans); P; P = P->next)
        if (P->n == n) return P;
    P = OS(make_plan)(n);
    P->next = OS(plans);
    OS(plans) = P;
    return P;
}

/* Bluestein p-point DFT, p = P->n prime: t (contiguous) -> y (contiguous). */
static void OS(bluestein)(const OS(plan) *P, int sign, const OS(cpx) *t,
                          long ts, OS(cpx) *y, long ys)
{
    long n = P->n, M = P->M, m;
    OS(cpx) *a = (OS(cpx) *)malloc(sizeof(OS(cpx)) * M);
    OS(cpx) *A = (OS(cpx) *)malloc(sizeof(OS(cpx)) * M);
    const OS(cpx) *B = sign < 0 ? P->bf : P->bb;
    for (m = 0; m < n; ++m) {
        OS(cpx) c = P->chirp[m];
        if (sign > 0) c.im = -c.im;
        a[m] = OS(cmul)(t[m * ts], c);
    }
    for (; m < M; ++m) a[m].re = a[m].im = 0;
    OS(fft)(P->sub, -1, a, 1, A);
    for (m = 0; m < M; ++m) A[m] = OS(cmul)(A[m], B[m]);
    OS(fft)(P->sub, +1, A, 1, a);
    for (m = 0; m < n; ++m) {
        OS(cpx) c = P->chirp[m];
        if (sign > 0) c.im = -c.im;
        y[m * ys] = OS(cmul)(a[m], c);
    }
    free(a);
    free(A);
}

/* Recursive DIT.  Sub-size nn = P->n / mul, factors from index fi. */
static void OS(rec)(const OS(plan) *P, int fi, long nn, long mul, int sign,
                    const OS(cpx) *x, long xs, OS(cpx) *y)
{
    long p, m, q, s, k;
    OS(cpx) *t, *o;
    if (nn == 1) { y[0] = x[0]; return; }
    p = P->f[fi];
    m = nn / p;
    if (m == 1 && p > ORC_DIRECT_MAX) {
        OS(bluestein)(OS(get_plan)(p), sign, x, xs, y, 1);
        return;
    }
    for (q = 0; q < p; ++q)
        OS(rec)(P, fi + 1, m, mul * p, sign, x + q * xs, xs * p, y + q * m);
    OS(cpx) sbuf[2 * (ORC_DIRECT_MAX + 1)];
    t = p <= ORC_DIRECT_MAX ? sbuf : (OS(cpx) *)malloc(sizeof(OS(cpx)) * 2 * p);
    o = t + p;
    for (k = 0; k < m; ++k) {
        for (q = 0; q < p; ++q)
            t[q] = OS(cmul)(y[q * m + k], OS(tw)(P, q * k, mul, sign));
        if (p > ORC_DIRECT_MAX) {
            OS(bluestein)(OS(get_plan)(p), sign, t, 1, o, 1);
        } else {
            for (s = 0; s < p; ++s) {
                OR re = 0, im = 0;
                for (q = 0; q < p; ++q) {
                    /* W_p^{q*s} = W_nn^{q*s*m} */
                    OS(cpx) w = OS(tw)(P, (q * s) % p * m, mul, sign);
                    re += t[q].re * w.re - t[q].im * w.im;
                    im += t[q].re * w.im + t[q].im * w.re;
                }
                o[s].re = re; o[s].im = im;
            }
        }
        for (s = 0; s < p; ++s) y[s * m + k] = o[s];
    }
    if (t != sbuf) free(t);
}

static void OS(fft)(const OS(plan) *P, int sign, const OS(cpx) *x, long xs,
                    OS(cpx) *y)
{
    OS(rec)(P, 0, P->n, 1, sign, x, xs, y);
}

/* Make sure every plan a length-n transform can reach exists. */
static void OS(prepare)(long n)
{
    OS(plan) *P = OS(get_plan)(n);
    int i;
    for (i = 0; i < P->nf; ++i)
        if (P->f[i] > ORC_DIRECT_MAX) OS(get_plan)(P->f[i]);
}

/* In-place separable multi-D DFT over H contiguous row-major blocks of
   shape n[0..rank-1]; dims [d0, d1) are transformed. */
static void OS(separable)(OS(cpx) *buf, long H, int rank, const long *n,
                          int d0, int d1, int sign)
{
    long total = 1;
    int d;
    for (d = 0; d < rank; ++d) total *= n[d];
    for (d = d0; d < d1; ++d) {
        long nd = n[d], inner = 1, lines, L;
        int e;
        OS(plan) *P;
        if (nd == 1) continue;
        OS(prepare)(nd);
        P = OS(get_plan)(nd);
        for (e = d + 1; e < rank; ++e) inner *= n[e];
        lines = H * (total / nd);
#pragma omp parallel
        {
            OS(cpx) *tin = (OS(cpx) *)malloc(sizeof(OS(cpx)) * nd);
            OS(cpx) *tout = (OS(cpx) *)malloc(sizeof(OS(cpx)) * nd);
#pragma omp for schedule(static)
            for (L = 0; L < lines; ++L) {
                long h = L / (total / nd), r = L % (total / nd);
                long outer = r / inner, in_ = r % inner, j;
                OS(cpx) *base = buf + h * total + outer * nd * inner + in_;
                for (j = 0; j < nd; ++j) tin[j] = base[j * inner];
                OS(fft)(P, sign, tin, 1, tout);
                for (j = 0; j < nd; ++j) base[j * inner] = tout[j];
            }
            free(tin);
            free(tout);
        }
    }
}
