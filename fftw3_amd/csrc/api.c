/*
 * fftw3.h API instance.  Compiled twice: once as fftw_ (double) and once
 * with -DFK_SINGLE as fftwf_ (float).
 *
 * Argument mapping follows the reference API layer:
 *   plan_many_dft      api/plan-many-dft.c:26-51 + api/mktensor-rowmajor.c:23-43
 *   plan_guru(64)_dft  api/plan-guru-dft.h:24-44 + api/mktensor-iodims.h:23-62
 *   plan_many_dft_r2c  api/plan-many-dft-r2c.c:24-57 (+ api/rdft2-pad.c:24-39)
 *   plan_many_dft_c2r  api/plan-many-dft-c2r.c:24-59
 *   plan_guru_dft_r2c/c2r api/plan-guru-dft-r2c.h:24-43, plan-guru-dft-c2r.h:24-45
 *   execute*           api/execute*.c
 *   copy/destroy_plan  api/apiplan.c:180-210
 * Strides stay in elements of their own array (complex or real), which is
 * FFTW's user-level convention; the x2 scaling FFTW applies internally is
 * not needed because the kernels address complex elements directly.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "engine.h"
#include "../../include/fftw3.h"
#include "../../include/fftw3_amd.h"

#ifdef FK_SINGLE
#define X(name) fftwf_##name
typedef float R;
typedef fftwf_complex C;
#define PREC 1
#else
#define X(name) fftw_##name
typedef double R;
typedef fftw_complex C;
#define PREC 0
#endif

typedef X(plan) PLAN;

static PLAN mk(fk_problem *pb)
{
    pb->prec = PREC;
    return (PLAN)fk_plan_create(pb);
}

static int many_kosher(int rank, const int *n, int howmany)
{
    int i;
    if (howmany < 0 || rank < 0 || rank > FK_MAXRANK) return 0;
    for (i = 0; i < rank; ++i)
        if (n[i] <= 0) return 0;
    return 1;
}

/* row-major tensor from (n, nembed, stride): api/mktensor-rowmajor.c:23-43 */
static void rowmajor(int rank, const int *n, const int *ni, const int *no, long long is,
                     long long os, fk_dim *d)
{
    int i;
    if (rank <= 0) return;
    d[rank - 1].n = n[rank - 1];
    d[rank - 1].is = is;
    d[rank - 1].os = os;
    for (i = rank - 1; i > 0; --i) {
        d[i - 1].n = n[i - 1];
        d[i - 1].is = d[i].is * ni[i];
        d[i - 1].os = d[i].os * no[i];
    }
}

/* ---------------- complex DFT ---------------- */

PLAN X(plan_many_dft)(int rank, const int *n, int howmany, C *in, const int *inembed,
                      int istride, int idist, C *out, const int *onembed, int ostride, int odist,
                      int sign, unsigned flags)
{
    fk_problem pb;
    if (!many_kosher(rank, n, howmany)) return 0;
    memset(&pb, 0, sizeof(pb));
    pb.kind = FK_C2C;
    pb.sign = sign == FFTW_FORWARD ? -1 : 1;
    pb.flags = flags;
    pb.rank = rank;
    rowmajor(rank, n, inembed ? inembed : n, onembed ? onembed : n, istride, ostride, pb.dims);
    pb.vrank = 1;
    pb.vdims[0].n = howmany;
    pb.vdims[0].is = idist;
    pb.vdims[0].os = odist;
    pb.in = in;
    pb.out = out;
    return mk(&pb);
}

PLAN X(plan_dft)(int rank, const int *n, C *in, C *out, int sign, unsigned flags)
{
    return X(plan_many_dft)(rank, n, 1, in, 0, 1, 1, out, 0, 1, 1, sign, flags);
}

PLAN X(plan_dft_1d)(int n, C *in, C *out, int sign, unsigned flags)
{
    return X(plan_dft)(1, &n, in, out, sign, flags);
}

PLAN X(plan_dft_2d)(int n0, int n1, C *in, C *out, int sign, unsigned flags)
{
    int n[2] = {n0, n1};
    return X(plan_dft)(2, n, in, out, sign, flags);
}

PLAN X(plan_dft_3d)(int n0, int n1, int n2, C *in, C *out, int sign, unsigned flags)
{
    int n[3] = {n0, n1, n2};
    return X(plan_dft)(3, n, in, out, sign, flags);
}

/* guru kosherness: api/mktensor-iodims.h:40-62 */
#define GURU_BODY(DIMT)                                                          \
    int i;                                                                       \
    if (rank < 0 || rank > FK_MAXRANK || howmany_rank < 0 ||                     \
        howmany_rank > FK_MAXRANK)                                               \
        return 0;                                                                \
    if (rank > 0 && !dims) return 0;                                             \
    if (howmany_rank > 0 && !howmany_dims) return 0;                             \
    for (i = 0; i < rank; ++i)                                                   \
        if (dims[i].n <= 0) return 0;                                            \
    for (i = 0; i < howmany_rank; ++i)                                           \
        if (howmany_dims[i].n < 0) return 0;                                     \
    memset(&pb, 0, sizeof(pb));                                                  \
    pb.flags = flags;                                                            \
    pb.rank = rank;                                                              \
    pb.vrank = howmany_rank;                                                     \
    for (i = 0; i < rank; ++i) {                                                 \
        pb.dims[i].n = dims[i].n;                                                \
        pb.dims[i].is = dims[i].is;                                              \
        pb.dims[i].os = dims[i].os;                                              \
    }                                                                            \
    for (i = 0; i < howmany_rank; ++i) {                                         \
        pb.vdims[i].n = howmany_dims[i].n;                                       \
        pb.vdims[i].is = howmany_dims[i].is;                                     \
        pb.vdims[i].os = howmany_dims[i].os;                                     \
    }

PLAN X(plan_guru_dft)(int rank, const X(iodim) *dims, int howmany_rank,
                      const X(iodim) *howmany_dims, C *in, C *out, int sign, unsigned flags)
{
    fk_problem pb;
    GURU_BODY(X(iodim))
    pb.kind = FK_C2C;
    pb.sign = sign == FFTW_FORWARD ? -1 : 1;
    pb.in = in;
    pb.out = out;
    return mk(&pb);
}

PLAN X(plan_guru64_dft)(int rank, const X(iodim64) *dims, int howmany_rank,
                        const X(iodim64) *howmany_dims, C *in, C *out, int sign, unsigned flags)
{
    fk_problem pb;
    GURU_BODY(X(iodim64))
    pb.kind = FK_C2C;
    pb.sign = sign == FFTW_FORWARD ? -1 : 1;
    pb.in = in;
    pb.out = out;
    return mk(&pb);
}

/* ---------------- split-complex arrays ----------------
   api/plan-guru-split-dft.c: real and imaginary parts in separate arrays,
   strides counting reals, no sign argument (forward on (ri, ii) -> (ro, io);
   the backward transform is obtained by swapping the pointers of each pair,
   doc/reference.texi "Guru Complex DFTs").  A pair that is one interleaved
   array (im = re + 1, or re = im + 1 for the swapped form, every stride
   even) runs on the interleaved kernels directly; separate arrays are
   staged through an interleaved buffer by the engine. */

/* side 0: input strides, 1: output strides */
static int split_pairing(const R *re, const R *im, const fk_problem *pb, int side)
{
    int i;
    for (i = 0; i < pb->rank; ++i)
        if ((side ? pb->dims[i].os : pb->dims[i].is) % 2) return FK_SPLIT_NONE;
    for (i = 0; i < pb->vrank; ++i)
        if ((side ? pb->vdims[i].os : pb->vdims[i].is) % 2) return FK_SPLIT_NONE;
    if (im == re + 1 && ((uintptr_t)re % sizeof(C)) == 0) return FK_SPLIT_PAIRED;
    if (re == im + 1 && ((uintptr_t)im % sizeof(C)) == 0) return FK_SPLIT_SWAPPED;
    return FK_SPLIT_NONE;
}

static void halve_strides(fk_problem *pb, int side)
{
    int i;
    for (i = 0; i < pb->rank; ++i) {
        if (side) pb->dims[i].os /= 2;
        else pb->dims[i].is /= 2;
    }
    for (i = 0; i < pb->vrank; ++i) {
        if (side) pb->vdims[i].os /= 2;
        else pb->vdims[i].is /= 2;
    }
}

static PLAN mk_split_c2c(fk_problem *pb, R *ri, R *ii, R *ro, R *io)
{
    int mi = split_pairing(ri, ii, pb, 0), mo = split_pairing(ro, io, pb, 1);
    pb->kind = FK_C2C;
    if (mi != FK_SPLIT_NONE && mi == mo) {
        fk_plan *P;
        halve_strides(pb, 0);
        halve_strides(pb, 1);
        pb->sign = mi == FK_SPLIT_PAIRED ? -1 : 1;
        pb->in = mi == FK_SPLIT_PAIRED ? (void *)ri : (void *)ii;
        pb->out = mo == FK_SPLIT_PAIRED ? (void *)ro : (void *)io;
        P = (fk_plan *)mk(pb);
        if (P) P->split = mi;
        return (PLAN)P;
    }
    pb->sign = -1;
    pb->split = FK_SPLIT_GENERAL;
    pb->in = ri;
    pb->in2 = ii;
    pb->out = ro;
    pb->out2 = io;
    return mk(pb);
}

PLAN X(plan_guru_split_dft)(int rank, const X(iodim) *dims, int howmany_rank,
                            const X(iodim) *howmany_dims, R *ri, R *ii, R *ro, R *io,
                            unsigned flags)
{
    fk_problem pb;
    GURU_BODY(X(iodim))
    return mk_split_c2c(&pb, ri, ii, ro, io);
}

PLAN X(plan_guru64_split_dft)(int rank, const X(iodim64) *dims, int howmany_rank,
                              const X(iodim64) *howmany_dims, R *ri, R *ii, R *ro, R *io,
                              unsigned flags)
{
    fk_problem pb;
    GURU_BODY(X(iodim64))
    return mk_split_c2c(&pb, ri, ii, ro, io);
}

static fk_plan *P_(PLAN p) { return (fk_plan *)p; }

/* api/execute.c:23-27: the planning arrays (both halves of split ones) */
void X(execute)(const PLAN p)
{
    fk_plan *P = P_(p);
    if (!P) return;
    if (fk_execute4(P, P->in, P->in2, P->out, P->out2)) {
        fprintf(stderr, "fftw3_amd: %s\n", fk_last_error());
        abort();
    }
}

/* new-array execute: api/execute-dft.c:25-32 (sign is part of the plan) */
void X(execute_dft)(const PLAN p, C *in, C *out)
{
    if (!p) return;
    fk_execute_or_die(P_(p), in, out);
}

static void die(const char *msg)
{
    fprintf(stderr, "fftw3_amd: %s\n", msg);
    abort();
}

static void exec4_or_die(fk_plan *P, void *in, void *in2, void *out, void *out2)
{
    if (fk_execute4(P, in, in2, out, out2)) die(fk_last_error());
}

/* api/execute-split-dft.c:25-29: new arrays of the same layout.  A plan
   that runs on interleaved kernels needs the same pairing as at planning. */
void X(execute_split_dft)(const PLAN p, R *ri, R *ii, R *ro, R *io)
{
    fk_plan *P = P_(p);
    if (!P) return;
    if (P->split == FK_SPLIT_GENERAL) {
        exec4_or_die(P, ri, ii, ro, io);
    } else if (P->split == FK_SPLIT_SWAPPED) {
        if (ri != ii + 1 || ro != io + 1)
            die("execute_split_dft: arrays must keep the pairing of the planning arrays");
        fk_execute_or_die(P, ii, io);
    } else {
        /* paired split plans, and interleaved plans given split views */
        if (ii != ri + 1 || io != ro + 1)
            die("execute_split_dft: arrays must keep the pairing of the planning arrays");
        fk_execute_or_die(P, ri, ro);
    }
}

/* ---------------- real DFT ---------------- */

/* api/rdft2-pad.c:24-39: default nembed of the padded (in-place) or
   complex side is n with the last dim n/2+1 complex / 2(n/2+1) real */
static const int *pad(int rank, const int *n, const int *nembed, int inplace, int cmplx,
                      int *buf)
{
    if (nembed || rank <= 0) return nembed ? nembed : n;
    if (inplace || cmplx) {
        memcpy(buf, n, sizeof(int) * rank);
        buf[rank - 1] = (n[rank - 1] / 2 + 1) * (1 + !cmplx);
        return buf;
    }
    return n;
}

PLAN X(plan_many_dft_r2c)(int rank, const int *n, int howmany, R *in, const int *inembed,
                          int istride, int idist, C *out, const int *onembed, int ostride,
                          int odist, unsigned flags)
{
    fk_problem pb;
    int bi[FK_MAXRANK], bo[FK_MAXRANK], inplace;
    if (!many_kosher(rank, n, howmany)) return 0;
    inplace = (void *)in == (void *)out;
    memset(&pb, 0, sizeof(pb));
    pb.kind = FK_R2C;
    pb.sign = -1;
    pb.flags = flags;
    pb.rank = rank;
    rowmajor(rank, n, pad(rank, n, inembed, inplace, 0, bi), pad(rank, n, onembed, inplace, 1, bo),
             istride, ostride, pb.dims);
    pb.vrank = 1;
    pb.vdims[0].n = howmany;
    pb.vdims[0].is = idist;
    pb.vdims[0].os = odist;
    pb.in = in;
    pb.out = out;
    return mk(&pb);
}

PLAN X(plan_dft_r2c)(int rank, const int *n, R *in, C *out, unsigned flags)
{
    return X(plan_many_dft_r2c)(rank, n, 1, in, 0, 1, 1, out, 0, 1, 1, flags);
}

PLAN X(plan_dft_r2c_1d)(int n, R *in, C *out, unsigned flags)
{
    return X(plan_dft_r2c)(1, &n, in, out, flags);
}

PLAN X(plan_dft_r2c_2d)(int n0, int n1, R *in, C *out, unsigned flags)
{
    int n[2] = {n0, n1};
    return X(plan_dft_r2c)(2, n, in, out, flags);
}

PLAN X(plan_dft_r2c_3d)(int n0, int n1, int n2, R *in, C *out, unsigned flags)
{
    int n[3] = {n0, n1, n2};
    return X(plan_dft_r2c)(3, n, in, out, flags);
}

PLAN X(plan_many_dft_c2r)(int rank, const int *n, int howmany, C *in, const int *inembed,
                          int istride, int idist, R *out, const int *onembed, int ostride,
                          int odist, unsigned flags)
{
    fk_problem pb;
    int bi[FK_MAXRANK], bo[FK_MAXRANK], inplace;
    if (!many_kosher(rank, n, howmany)) return 0;
    inplace = (void *)in == (void *)out;
    if (!inplace) flags |= FFTW_DESTROY_INPUT;
    memset(&pb, 0, sizeof(pb));
    pb.kind = FK_C2R;
    pb.sign = 1;
    pb.flags = flags;
    pb.rank = rank;
    rowmajor(rank, n, pad(rank, n, inembed, inplace, 1, bi), pad(rank, n, onembed, inplace, 0, bo),
             istride, ostride, pb.dims);
    pb.vrank = 1;
    pb.vdims[0].n = howmany;
    pb.vdims[0].is = idist;
    pb.vdims[0].os = odist;
    pb.in = in;
    pb.out = out;
    return mk(&pb);
}

PLAN X(plan_dft_c2r)(int rank, const int *n, C *in, R *out, unsigned flags)
{
    return X(plan_many_dft_c2r)(rank, n, 1, in, 0, 1, 1, out, 0, 1, 1, flags);
}

PLAN X(plan_dft_c2r_1d)(int n, C *in, R *out, unsigned flags)
{
    return X(plan_dft_c2r)(1, &n, in, out, flags);
}

PLAN X(plan_dft_c2r_2d)(int n0, int n1, C *in, R *out, unsigned flags)
{
    int n[2] = {n0, n1};
    return X(plan_dft_c2r)(2, n, in, out, flags);
}

PLAN X(plan_dft_c2r_3d)(int n0, int n1, int n2, C *in, R *out, unsigned flags)
{
    int n[3] = {n0, n1, n2};
    return X(plan_dft_c2r)(3, n, in, out, flags);
}

PLAN X(plan_guru_dft_r2c)(int rank, const X(iodim) *dims, int howmany_rank,
                          const X(iodim) *howmany_dims, R *in, C *out, unsigned flags)
{
    fk_problem pb;
    GURU_BODY(X(iodim))
    pb.kind = FK_R2C;
    pb.sign = -1;
    pb.in = in;
    pb.out = out;
    return mk(&pb);
}

PLAN X(plan_guru64_dft_r2c)(int rank, const X(iodim64) *dims, int howmany_rank,
                            const X(iodim64) *howmany_dims, R *in, C *out, unsigned flags)
{
    fk_problem pb;
    GURU_BODY(X(iodim64))
    pb.kind = FK_R2C;
    pb.sign = -1;
    pb.in = in;
    pb.out = out;
    return mk(&pb);
}

PLAN X(plan_guru_dft_c2r)(int rank, const X(iodim) *dims, int howmany_rank,
                          const X(iodim) *howmany_dims, C *in, R *out, unsigned flags)
{
    fk_problem pb;
    if ((void *)out != (void *)in) flags |= FFTW_DESTROY_INPUT;
    {
        GURU_BODY(X(iodim))
        pb.kind = FK_C2R;
        pb.sign = 1;
        pb.in = in;
        pb.out = out;
    }
    return mk(&pb);
}

PLAN X(plan_guru64_dft_c2r)(int rank, const X(iodim64) *dims, int howmany_rank,
                            const X(iodim64) *howmany_dims, C *in, R *out, unsigned flags)
{
    fk_problem pb;
    if ((void *)out != (void *)in) flags |= FFTW_DESTROY_INPUT;
    {
        GURU_BODY(X(iodim64))
        pb.kind = FK_C2R;
        pb.sign = 1;
        pb.in = in;
        pb.out = out;
    }
    return mk(&pb);
}

static PLAN mk_split_r2c(fk_problem *pb, R *in, R *ro, R *io)
{
    pb->kind = FK_R2C;
    pb->sign = -1;
    pb->in = in;
    if (split_pairing(ro, io, pb, 1) == FK_SPLIT_PAIRED) {
        fk_plan *P;
        halve_strides(pb, 1);
        pb->out = ro;
        P = (fk_plan *)mk(pb);
        if (P) P->split = FK_SPLIT_PAIRED;
        return (PLAN)P;
    }
    pb->split = FK_SPLIT_GENERAL;
    pb->out = ro;
    pb->out2 = io;
    return mk(pb);
}

static PLAN mk_split_c2r(fk_problem *pb, R *ri, R *ii, R *out)
{
    pb->kind = FK_C2R;
    pb->sign = 1;
    pb->out = out;
    if (split_pairing(ri, ii, pb, 0) == FK_SPLIT_PAIRED) {
        fk_plan *P;
        halve_strides(pb, 0);
        pb->in = ri;
        if ((void *)ri != (void *)out) pb->flags |= FFTW_DESTROY_INPUT;
        P = (fk_plan *)mk(pb);
        if (P) P->split = FK_SPLIT_PAIRED;
        return (PLAN)P;
    }
    /* staged: the caller's arrays are only read */
    pb->split = FK_SPLIT_GENERAL;
    pb->in = ri;
    pb->in2 = ii;
    return mk(pb);
}

PLAN X(plan_guru_split_dft_r2c)(int rank, const X(iodim) *dims, int howmany_rank,
                                const X(iodim) *howmany_dims, R *in, R *ro, R *io,
                                unsigned flags)
{
    fk_problem pb;
    GURU_BODY(X(iodim))
    return mk_split_r2c(&pb, in, ro, io);
}

PLAN X(plan_guru_split_dft_c2r)(int rank, const X(iodim) *dims, int howmany_rank,
                                const X(iodim) *howmany_dims, R *ri, R *ii, R *out,
                                unsigned flags)
{
    fk_problem pb;
    GURU_BODY(X(iodim))
    return mk_split_c2r(&pb, ri, ii, out);
}

PLAN X(plan_guru64_split_dft_r2c)(int rank, const X(iodim64) *dims, int howmany_rank,
                                  const X(iodim64) *howmany_dims, R *in, R *ro, R *io,
                                  unsigned flags)
{
    fk_problem pb;
    GURU_BODY(X(iodim64))
    return mk_split_r2c(&pb, in, ro, io);
}

PLAN X(plan_guru64_split_dft_c2r)(int rank, const X(iodim64) *dims, int howmany_rank,
                                  const X(iodim64) *howmany_dims, R *ri, R *ii, R *out,
                                  unsigned flags)
{
    fk_problem pb;
    GURU_BODY(X(iodim64))
    return mk_split_c2r(&pb, ri, ii, out);
}

/* api/execute-dft-r2c.c:25-30 / execute-dft-c2r.c:25-30 */
void X(execute_dft_r2c)(const PLAN p, R *in, C *out)
{
    if (!p) return;
    fk_execute_or_die(P_(p), in, out);
}

void X(execute_dft_c2r)(const PLAN p, C *in, R *out)
{
    if (!p) return;
    fk_execute_or_die(P_(p), in, out);
}

void X(execute_split_dft_r2c)(const PLAN p, R *in, R *ro, R *io)
{
    fk_plan *P = P_(p);
    if (!P) return;
    if (P->split == FK_SPLIT_GENERAL) {
        exec4_or_die(P, in, 0, ro, io);
    } else {
        if (io != ro + 1)
            die("execute_split_dft_r2c: arrays must keep the pairing of the planning arrays");
        fk_execute_or_die(P, in, ro);
    }
}

void X(execute_split_dft_c2r)(const PLAN p, R *ri, R *ii, R *out)
{
    fk_plan *P = P_(p);
    if (!P) return;
    if (P->split == FK_SPLIT_GENERAL) {
        exec4_or_die(P, ri, ii, out, 0);
    } else {
        if (ii != ri + 1)
            die("execute_split_dft_c2r: arrays must keep the pairing of the planning arrays");
        fk_execute_or_die(P, ri, out);
    }
}

/* ---------------- real-to-real kinds ----------------
   api/plan-many-r2r.c:26-51, api/plan-guru-r2r.h:23-44, api/plan-r2r.c,
   api/plan-r2r-{1,2,3}d.c; strides and dists count reals.  The kind of
   each dim maps one to one (api/map-r2r-kind.c:24-50); an unknown kind
   makes the planner return NULL. */

static int r2r_kinds(fk_problem *pb, int rank, const X(r2r_kind) *kind)
{
    int i;
    if (rank > 0 && !kind) return 0;
    for (i = 0; i < rank; ++i) {
        if ((int)kind[i] < FFTW_R2HC || (int)kind[i] > FFTW_RODFT11) return 0;
        pb->r2r[i] = (int)kind[i];
    }
    return 1;
}

PLAN X(plan_many_r2r)(int rank, const int *n, int howmany, R *in, const int *inembed,
                      int istride, int idist, R *out, const int *onembed, int ostride, int odist,
                      const X(r2r_kind) *kind, unsigned flags)
{
    fk_problem pb;
    if (!many_kosher(rank, n, howmany)) return 0;
    memset(&pb, 0, sizeof(pb));
    if (!r2r_kinds(&pb, rank, kind)) return 0;
    pb.kind = FK_R2R;
    pb.flags = flags;
    pb.rank = rank;
    rowmajor(rank, n, inembed ? inembed : n, onembed ? onembed : n, istride, ostride, pb.dims);
    pb.vrank = 1;
    pb.vdims[0].n = howmany;
    pb.vdims[0].is = idist;
    pb.vdims[0].os = odist;
    pb.in = in;
    pb.out = out;
    return mk(&pb);
}

PLAN X(plan_r2r)(int rank, const int *n, R *in, R *out, const X(r2r_kind) *kind,
                 unsigned flags)
{
    return X(plan_many_r2r)(rank, n, 1, in, 0, 1, 1, out, 0, 1, 1, kind, flags);
}

PLAN X(plan_r2r_1d)(int n, R *in, R *out, X(r2r_kind) kind, unsigned flags)
{
    return X(plan_r2r)(1, &n, in, out, &kind, flags);
}

PLAN X(plan_r2r_2d)(int n0, int n1, R *in, R *out, X(r2r_kind) kind0, X(r2r_kind) kind1,
                    unsigned flags)
{
    int n[2] = {n0, n1};
    X(r2r_kind) k[2] = {kind0, kind1};
    return X(plan_r2r)(2, n, in, out, k, flags);
}

PLAN X(plan_r2r_3d)(int n0, int n1, int n2, R *in, R *out, X(r2r_kind) kind0,
                    X(r2r_kind) kind1, X(r2r_kind) kind2, unsigned flags)
{
    int n[3] = {n0, n1, n2};
    X(r2r_kind) k[3] = {kind0, kind1, kind2};
    return X(plan_r2r)(3, n, in, out, k, flags);
}

PLAN X(plan_guru_r2r)(int rank, const X(iodim) *dims, int howmany_rank,
                      const X(iodim) *howmany_dims, R *in, R *out, const X(r2r_kind) *kind,
                      unsigned flags)
{
    fk_problem pb;
    GURU_BODY(X(iodim))
    if (!r2r_kinds(&pb, rank, kind)) return 0;
    pb.kind = FK_R2R;
    pb.in = in;
    pb.out = out;
    return mk(&pb);
}

PLAN X(plan_guru64_r2r)(int rank, const X(iodim64) *dims, int howmany_rank,
                        const X(iodim64) *howmany_dims, R *in, R *out, const X(r2r_kind) *kind,
                        unsigned flags)
{
    fk_problem pb;
    GURU_BODY(X(iodim64))
    if (!r2r_kinds(&pb, rank, kind)) return 0;
    pb.kind = FK_R2R;
    pb.in = in;
    pb.out = out;
    return mk(&pb);
}

/* api/execute-r2r.c:24-28: new arrays, same layout and alignment */
void X(execute_r2r)(const PLAN p, R *in, R *out)
{
    if (!p) return;
    fk_execute_or_die(P_(p), in, out);
}

/* ---------------- plan lifetime (api/apiplan.c:180-210) ---------------- */

PLAN X(copy_plan)(PLAN p)
{
    if (p) __sync_fetch_and_add(&P_(p)->refcount, 1);
    return p;
}

void X(destroy_plan)(PLAN p)
{
    if (p && __sync_sub_and_fetch(&P_(p)->refcount, 1) == 0) fk_plan_destroy(P_(p));
}

void X(forget_wisdom)(void) { fk_forget_wisdom(); }
void X(cleanup)(void) { fk_cleanup(); }
void X(set_timelimit)(double t) { fk_set_timelimit(t); }

/* threads API: the device does the parallel work; these keep their
   contracts (init succeeds, thread count recorded) */
static int g_nthreads = 1;
void X(plan_with_nthreads)(int nthreads) { g_nthreads = nthreads < 1 ? 1 : nthreads; }
int X(planner_nthreads)(void) { return g_nthreads; }
int X(init_threads)(void) { return 1; }
void X(cleanup_threads)(void) {}
void X(threads_set_callback)(void (*parallel_loop)(void *(*work)(char *), char *jobdata,
                                                   size_t elsize, int njobs, void *data),
                             void *data)
{
    (void)parallel_loop;
    (void)data;
}
void X(make_planner_thread_safe)(void) {}

/* ---------------- wisdom ---------------- */

char *X(export_wisdom_to_string)(void) { return fk_wisdom_export(); }

void X(export_wisdom)(X(write_char_func) write_char, void *data)
{
    char *s = fk_wisdom_export(), *q;
    for (q = s; q && *q; ++q) write_char(*q, data);
    free(s);
}

void X(export_wisdom_to_file)(FILE *f)
{
    char *s = fk_wisdom_export();
    if (s) fputs(s, f);
    free(s);
}

int X(export_wisdom_to_filename)(const char *filename)
{
    FILE *f = fopen(filename, "w");
    int ok;
    if (!f) return 0;
    X(export_wisdom_to_file)(f);
    ok = !ferror(f);
    return fclose(f) == 0 && ok;
}

int X(import_wisdom_from_string)(const char *s) { return s ? fk_wisdom_import(s) : 0; }

int X(import_wisdom)(X(read_char_func) read_char, void *data)
{
    size_t cap = 4096, len = 0;
    char *s = (char *)malloc(cap);
    int ch, r;
    while ((ch = read_char(data)) != EOF) {
        if (len + 2 > cap) s = (char *)realloc(s, cap *= 2);
        s[len++] = (char)ch;
    }
    s[len] = 0;
    r = fk_wisdom_import(s);
    free(s);
    return r;
}

static int rd_file(void *d) { return fgetc((FILE *)d); }

int X(import_wisdom_from_file)(FILE *f) { return X(import_wisdom)(rd_file, f); }

int X(import_wisdom_from_filename)(const char *filename)
{
    FILE *f = fopen(filename, "r");
    int r;
    if (!f) return 0;
    r = X(import_wisdom_from_file)(f);
    fclose(f);
    return r;
}

int X(import_system_wisdom)(void)
{
    const char *p = getenv("FFTW_AMD_SYSTEM_WISDOM");
    return X(import_wisdom_from_filename)(p ? p : "/etc/fftw/wisdom-gfx950");
}

/* ---------------- printing ---------------- */

void X(fprint_plan)(const PLAN p, FILE *f) { fk_plan_print(P_(p), f); }
void X(print_plan)(const PLAN p) { fk_plan_print(P_(p), stdout); }
char *X(sprint_plan)(const PLAN p) { return fk_plan_sprint(P_(p)); }

/* ---------------- memory: device allocations ---------------- */

void *X(malloc)(size_t n) { return fk_malloc(n); }
R *X(alloc_real)(size_t n) { return (R *)fk_malloc(sizeof(R) * n); }
C *X(alloc_complex)(size_t n) { return (C *)fk_malloc(sizeof(C) * n); }
void X(free)(void *p) { fk_free(p); }

/* ---------------- costs ---------------- */

void X(flops)(const PLAN p, double *add, double *mul, double *fmas)
{
    *add = p ? P_(p)->adds : 0;
    *mul = p ? P_(p)->muls : 0;
    *fmas = p ? P_(p)->fmas : 0;
}

/* modelled seconds per execute, from the plan's passes (api/flops.c:34-40
   returns the planner's cost estimate) */
double X(estimate_cost)(const PLAN p) { return p ? fk_plan_estimate(P_(p)) : 0; }

double X(cost)(const PLAN p) { return p ? P_(p)->pcost : 0; }

int X(alignment_of)(R *p) { return (int)((uintptr_t)p % 16); }

#ifdef FK_SINGLE
const char fftwf_version[] = "fftw-3.3-compatible fftw3-amd-gfx950-0.1";
const char fftwf_cc[] = "hipcc --offload-arch=gfx950";
const char fftwf_codelet_optim[] = "stockham-lds";
#else
const char fftw_version[] = "fftw-3.3-compatible fftw3-amd-gfx950-0.1";
const char fftw_cc[] = "hipcc --offload-arch=gfx950";
const char fftw_codelet_optim[] = "stockham-lds";
#endif

/* ---------------- extensions (defined once, in the double instance) ---- */
#ifndef FK_SINGLE
int fftw_amd_init(void) { return fk_init(); }
const char *fftw_amd_last_error(void) { return fk_last_error(); }
void fftw_amd_set_stream(void *s) { fk_set_stream((hipStream_t)s); }
void *fftw_amd_get_stream(void) { return (void *)fk_get_stream(); }
void fftw_amd_plan_set_stream(void *p, void *s) { fk_plan_set_stream((fk_plan *)p, (hipStream_t)s); }
void fftw_amd_set_async(int on) { fk_set_async(on); }

int fftw_amd_memcpy_htod(void *d, const void *s, size_t b)
{
    return (int)hipMemcpy(d, s, b, hipMemcpyHostToDevice);
}
int fftw_amd_memcpy_dtoh(void *d, const void *s, size_t b)
{
    return (int)hipMemcpy(d, s, b, hipMemcpyDeviceToHost);
}
int fftw_amd_memcpy_dtod(void *d, const void *s, size_t b)
{
    return (int)hipMemcpy(d, s, b, hipMemcpyDeviceToDevice);
}
int fftw_amd_memset(void *d, int v, size_t b) { return (int)hipMemset(d, v, b); }
int fftw_amd_stream_copy(void *d, const void *s, size_t b, int mode, int ctas, void *stream)
{
    return fkm_stream_copy(d, s, (long long)b, mode, ctas, (hipStream_t)stream);
}
int fftw_amd_synchronize(void) { return (int)hipDeviceSynchronize(); }

size_t fftw_amd_mem_free(void)
{
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return 0;
    return f;
}

int fftw_amd_fill_uniform(void *dev, int prec, size_t count, uint64_t seed, uint64_t stream)
{
    int rc = fkm_fill(dev, prec, (long long)count, seed, stream, 0, fk_get_stream());
    return rc ? rc : (int)hipStreamSynchronize(fk_get_stream());
}

int fftw_amd_fill_gauss(void *dev, int prec, size_t count, uint64_t seed, uint64_t stream)
{
    int rc = fkm_fill(dev, prec, (long long)count, seed, stream, 1, fk_get_stream());
    return rc ? rc : (int)hipStreamSynchronize(fk_get_stream());
}

int fftw_amd_plan_npasses(const void *p) { return fk_plan_npasses((const fk_plan *)p); }
int fftw_amd_plan_pass_info(const void *p, int i, char *buf, int len)
{
    return fk_plan_pass_info((const fk_plan *)p, i, buf, len);
}
double fftw_amd_plan_bytes(const void *p) { return fk_plan_bytes((const fk_plan *)p); }
double fftw_amd_plan_pass_bytes(const void *p, int i)
{
    const fk_plan *q = (const fk_plan *)p;
    return q && i >= 0 && i < q->npass ? q->pass[i].bytes : -1;
}
double fftw_amd_plan_time(void *p, void *in, void *out, int iters, double *per_pass_ms)
{
    return p ? fk_plan_time((fk_plan *)p, in, out, iters, per_pass_ms) : -1;
}

/* The new-array execute (execute_dft / _r2c / _c2r / _r2r) with the
   library's static device scratch in place of the plan's own: allocates
   nothing.  Returns 0 or a HIP error code (no abort). */
int fftw_amd_execute_noalloc(const void *p, void *in, void *out)
{
    return fk_execute_noalloc((fk_plan *)p, in, out);
}

int fftw_amd_catalog_size(void) { return fk_catalog_size(); }
int fftw_amd_catalog_entry(int i, int *info)
{
    const fk_kernel *k = fk_catalog_entry(i);
    int j;
    if (!k || !info) return -1;
    info[0] = k->prec;
    info[1] = k->L;
    info[2] = k->mode;
    info[3] = k->io;
    info[4] = k->tpl;
    info[5] = k->nl;
    info[6] = k->variant;
    info[7] = k->nrad;
    for (j = 0; j < 12; ++j) info[8 + j] = j < k->nrad ? k->rad[j] : 0;
    info[20] = (int)k->lds;
    info[21] = k->loop;
    return 0;
}
int fftw_amd_force_kernel(int i) { return fk_force_kernel(i); }

int fftw_amd_plan_use_mfma(void *p) { return fk_plan_use_mfma((fk_plan *)p); }
int fftw_amd_mfma_probe(float *dev_out)
{
    int rc = fkm_mfma_probe(dev_out, 0);
    return rc ? rc : (int)hipDeviceSynchronize();
}

fftw_amd_batch fftw_amd_plan_batch_dft_1d(int prec, int ngroups, const int *n,
                                          const int *howmany, void *const *in, void *const *out,
                                          int sign, unsigned flags)
{
    return (fftw_amd_batch)fk_batch_create(prec, ngroups, n, howmany, in, out, sign, flags);
}
int fftw_amd_batch_execute(fftw_amd_batch b) { return fk_batch_execute((fk_batch *)b, 1); }
int fftw_amd_batch_execute_direct(fftw_amd_batch b) { return fk_batch_execute((fk_batch *)b, 0); }
int fftw_amd_batch_info(fftw_amd_batch b, char *buf, int len)
{
    return fk_batch_info((fk_batch *)b, buf, len);
}
void *fftw_amd_batch_stream(fftw_amd_batch b) { return (void *)fk_batch_stream((fk_batch *)b); }
void fftw_amd_batch_destroy(fftw_amd_batch b) { fk_batch_destroy((fk_batch *)b); }

fftw_amd_server fftw_amd_server_create(size_t budget, int max_rounds, int per_round)
{
    return (fftw_amd_server)fk_server_create(budget, max_rounds, per_round);
}
int fftw_amd_server_prepare(fftw_amd_server s, const fftw_amd_request *r)
{
    return fk_server_prepare((fk_server *)s, (const fk_request *)r);
}
long long fftw_amd_server_submit(fftw_amd_server s, const fftw_amd_request *r)
{
    return fk_server_submit((fk_server *)s, (const fk_request *)r);
}
int fftw_amd_server_poll(fftw_amd_server s, long long id) { return fk_server_poll((fk_server *)s, id); }
int fftw_amd_server_wait(fftw_amd_server s, long long id, void **out, double *lat)
{
    return fk_server_wait((fk_server *)s, id, out, lat);
}
int fftw_amd_server_release(fftw_amd_server s, long long id)
{
    return fk_server_release((fk_server *)s, id);
}
int fftw_amd_server_stats(fftw_amd_server s, double *p50, double *p99, long long *done,
                          long long *transforms, double *elapsed, size_t *inuse)
{
    return fk_server_stats((fk_server *)s, p50, p99, done, transforms, elapsed, inuse);
}
void fftw_amd_server_destroy(fftw_amd_server s) { fk_server_destroy((fk_server *)s); }

void *fftw_amd_event_create(void)
{
    hipEvent_t e = 0;
    if (hipEventCreate(&e) != hipSuccess) return 0;
    return (void *)e;
}
void fftw_amd_event_record(void *ev, void *s) { (void)hipEventRecord((hipEvent_t)ev, (hipStream_t)s); }
double fftw_amd_event_elapsed(void *a, void *b)
{
    float ms = -1;
    if (hipEventSynchronize((hipEvent_t)b) != hipSuccess) return -1;
    if (hipEventElapsedTime(&ms, (hipEvent_t)a, (hipEvent_t)b) != hipSuccess) return -1;
    return ms;
}
void fftw_amd_event_destroy(void *ev) { (void)hipEventDestroy((hipEvent_t)ev); }
#endif
