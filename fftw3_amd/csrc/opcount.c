/*
 * Operation counts of a plan: the additions, multiplications and fused
 * multiply-adds its passes execute, for fftw_flops (api/flops.c:23-31).
 * FFTW sums the counts its solvers record while planning (kernel/ops.c:24-60,
 * opcnt {add, mul, fma, other}); here every pass is a kernel of known
 * structure, so the counts are derived from that structure:
 *
 *   - a Stockham pass runs, per line and stage s of radix R, L/R butterflies,
 *     each one compile-time DFT_R (fk_device.h Dft<>: 2- and 4-point forms,
 *     the symmetric pair form of odd primes, A x B splits with inter-factor
 *     twiddles whose multiples of pi/4 are specialised) plus, for s > 0,
 *     R - 1 complex twiddle products;
 *   - the IO kind adds its own arithmetic around the stages (the r2c
 *     post-pass, the c2r pre-pass, the Bluestein chirp and kernel products,
 *     the real-to-real reorderings and output twiddles), as do the output
 *     post-ops (four-step twiddle, scale);
 *   - element-wise passes count the arithmetic of their op per element.
 *
 * Counting convention: the kernels are compiled with floating-point
 * contraction, so a product feeding an addition (a*b + c, c - a*b) counts as
 * one fma, a complex product as 2 mul + 2 fma, and sign flips, swaps and
 * copies count nothing.  fftw_flops' total is add + mul + 2 fma.
 */
#include <string.h>
#include "engine.h"

typedef struct { double add, mul, fma; } ops;

static void ops_acc(ops *o, double s, const ops *a)
{
    o->add += s * a->add;
    o->mul += s * a->mul;
    o->fma += s * a->fma;
}

/* mirror of first_factor() in fk_device.h: how Dft<N> splits N = A * B */
static int first_factor(int n)
{
    int d;
    if (n % 4 == 0 && n != 4) {
        int a = 4;
        while (a * a < n && n % (a * 2) == 0 && (n / (a * 2)) >= 2) a *= 2;
        if (n % a == 0) return a;
        return 4;
    }
    for (d = 2; d * d <= n; ++d)
        if (n % d == 0) return d;
    return n;
}

/* v *= W_N^e (fk_device.h twc<>): free at multiples of pi/2, 2 add + 2 mul
   at odd multiples of pi/4, a general complex product otherwise */
static void twc_ops(int N, int e, ops *o)
{
    e = ((e % N) + N) % N;
    if (e == 0 || 4 * e == N || 2 * e == N || 4 * e == 3 * N) return;
    if (8 * e == N || 8 * e == 3 * N || 8 * e == 5 * N || 8 * e == 7 * N) {
        o->add += 2;
        o->mul += 2;
        return;
    }
    o->mul += 2;
    o->fma += 2;
}

/* one forward DFT of compile-time length N (fk_device.h Dft<R, N, S>) */
void fk_dft_ops(int N, double *add, double *mul, double *fma)
{
    ops o = {0, 0, 0};
    if (N == 2) {
        o.add = 4;
    } else if (N == 4) {
        o.add = 16;
    } else if (N >= 3) {
        int A = first_factor(N);
        if (A == N) {
            /* odd prime, H pairs: sums and differences (4H), DC (2H), and
               per output pair 4 accumulations of H terms plus 4 adds */
            double H = (N - 1) / 2;
            o.add = 10 * H;
            o.fma = 4 * H * H;
        } else {
            int B = N / A, k1, n2;
            double a1, m1, f1, a2, m2, f2;
            fk_dft_ops(A, &a1, &m1, &f1);
            fk_dft_ops(B, &a2, &m2, &f2);
            o.add = B * a1 + A * a2;
            o.mul = B * m1 + A * m2;
            o.fma = B * f1 + A * f2;
            for (k1 = 1; k1 < A; ++k1)
                for (n2 = 1; n2 < B; ++n2) twc_ops(N, k1 * n2, &o);
        }
    }
    *add = o.add;
    *mul = o.mul;
    *fma = o.fma;
}

static const ops CMUL = {0, 2, 2};

/* the stages of one line of a catalog kernel (twiddles on stages s > 0) */
static void stages_ops(const fk_kernel *k, ops *o)
{
    int s, ns = 1;
    for (s = 0; s < k->nrad; ++s) {
        const int R = k->rad[s];
        const double nb = (double)k->L / R;
        ops b;
        fk_dft_ops(R, &b.add, &b.mul, &b.fma);
        ops_acc(o, nb, &b);
        if (s > 0) ops_acc(o, nb * (R - 1), &CMUL);
        ns *= R;
    }
}

/* split-precision matrix-core radix-16 pass (fk_mfma.hip): per butterfly,
   three 32x32 real matrix products of the 16-point DFT (hi*hi, hi*lo,
   lo*hi) on the matrix cores, plus the stage twiddles */
static void mfma_ops(const fk_kernel *k, ops *o)
{
    int s;
    for (s = 0; s < k->nrad; ++s) {
        const double nb = (double)k->L / 16;
        o->fma += nb * 3.0 * 32 * 32;
        o->mul += nb * 32;             /* scale back */
        if (s > 0) ops_acc(o, nb * 15, &CMUL);
    }
}

/* one Stockham pass over all its lines */
static void stockham_ops(const fk_pass *ps, ops *o)
{
    const fk_kernel *k = ps->k;
    const double L = k->L, lines = (double)ps->a.nlines;
    const unsigned fl = ps->a.flags;
    ops t = {0, 0, 0};
    double outs = L;    /* elements that get the output post-ops */
    if (k->variant >= 100) mfma_ops(k, &t);
    else stages_ops(k, &t);
    switch (k->io) {
    case FK_IO_R2C:
    case FK_IO_R2C_RR:
    case FK_IO_R2C_XR:
        /* L+1 bins: E and O (4 add, 4 mul by 1/2), X = E + W O (4 fma) */
        outs = L + 1;
        t.add += 4 * outs;
        t.mul += 4 * outs;
        t.fma += 4 * outs;
        if (k->io == FK_IO_R2C_XR) {
            /* per bin and row: 1/nl of a DFT_nl across the tile's rows and
               (nl-1)/nl of a twiddle product */
            double da, dm, df;
            fk_dft_ops(k->nl, &da, &dm, &df);
            t.add += outs * da / k->nl;
            t.mul += outs * dm / k->nl;
            t.fma += outs * df / k->nl;
            ops_acc(&t, outs * (k->nl - 1) / k->nl, &CMUL);
        }
        if (k->io == FK_IO_R2C_RR) {
            const unsigned rm = ps->a.rmode;
            if (rm == FK_RM_DHT) t.add += 2 * outs;
            if (rm == FK_RM_DCT2 || rm == FK_RM_DST2) {
                ops_acc(&t, outs, &CMUL);
                t.mul += 2 * outs;
            }
        }
        break;
    case FK_IO_C2R:
    case FK_IO_C2R_RR:
    case FK_IO_C2R_XR:
        /* pre-pass per complex input: E, D (4 add), conj-twiddle product
           (2 mul + 2 fma), Z = E + i O (2 add) */
        t.add += 6 * L;
        t.mul += 2 * L;
        t.fma += 2 * L;
        if (k->io == FK_IO_C2R_XR) {
            /* per bin and row: the cross-row inverse stage, as in R2C_XR */
            double da, dm, df;
            fk_dft_ops(k->nl, &da, &dm, &df);
            t.add += (L + 1) * da / k->nl;
            t.mul += (L + 1) * dm / k->nl;
            t.fma += (L + 1) * df / k->nl;
            ops_acc(&t, (L + 1) * (k->nl - 1) / k->nl, &CMUL);
        }
        if (k->io == FK_IO_C2R_RR &&
            (ps->a.rmode == FK_RM_DCT3 || ps->a.rmode == FK_RM_DST3)) {
            /* exp(+i pi k / 2n)(x_k - i x_{n-k}) per bin */
            t.mul += 2 * (L + 1);
            t.fma += 2 * (L + 1);
        }
        break;
    case FK_IO_BLUE: {
        /* chirp in (n), kernel product (L), second L-point DFT, chirp out (n) */
        const double n = (double)ps->a.nvalid;
        ops_acc(&t, 1.0, &t);            /* forward and inverse DFT_L */
        ops_acc(&t, 2 * n + L, &CMUL);
        outs = n;
        break;
    }
    case FK_IO_DCT4:
        /* input twiddle (L) and output twiddle (L) products, 2 L scalings */
        ops_acc(&t, 2 * L, &CMUL);
        t.mul += 2 * L;
        break;
    case FK_IO_R2C_FULL:
        outs = (double)(k->L / 2 + 1);
        break;
    default:
        break;
    }
    if (fl & FK_TW_OUT) {
        /* FP64 twiddle (table product or recurrence step) and the product
           with the output, per element */
        ops_acc(&t, 2 * L, &CMUL);
    }
    if (fl & FK_SCALE) t.mul += 2 * outs;
    ops_acc(o, lines, &t);
}

/* arithmetic of one element of an element-wise pass */
static void ew_ops(const fk_pass *ps, ops *o)
{
    const fk_ew_args *e = &ps->e;
    const double cnt = (double)e->n * (double)e->nlines;
    double valid = cnt;
    ops t = {0, 0, 0};
    switch (e->op) {
    case FK_EW_BS_PREP:
        valid = (double)e->p0 * (double)e->nlines;
        /* fallthrough */
    case FK_EW_BS_MUL:
    case FK_EW_BS_POST:
    case FK_EW_DCT4_PRE:
    case FK_EW_DCT3_PRE:
        t = CMUL;
        break;
    case FK_EW_R2C_POST:
        t.add = 4;
        t.mul = 4;
        t.fma = 4;
        break;
    case FK_EW_C2R_PRE:
        t.add = 6;
        t.mul = 2;
        t.fma = 2;
        break;
    case FK_EW_HC_POST:
        if (e->p1) t.add = 1;
        break;
    case FK_EW_DCT2_POST:
        t.mul = 2;
        t.fma = 1;
        break;
    case FK_EW_DCT4_POST:
    case FK_EW_DCT4O_POST:
        t = CMUL;
        t.mul += 1;
        break;
    case FK_EW_DCT4O_PRE:
        valid = (double)e->p0 * (double)e->nlines;
        t.mul = 2;
        break;
    default:
        break;
    }
    if (e->flags & FK_SCALE) t.mul += (e->op == FK_EW_COPY_R || e->op == FK_EW_REAL) ? 1 : 2;
    ops_acc(o, valid, &t);
}

void fk_count_ops(fk_plan *P)
{
    ops o = {0, 0, 0};
    int i;
    for (i = 0; i < P->npass; ++i) {
        const fk_pass *ps = P->pass + i;
        if (ps->type == FK_PASS_STOCKHAM) stockham_ops(ps, &o);
        else ew_ops(ps, &o);
    }
    P->adds = o.add;
    P->muls = o.mul;
    P->fmas = o.fma;
}
