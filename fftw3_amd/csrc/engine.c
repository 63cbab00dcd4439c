/*
 * Host engine: problem -> pass list, device tables, execution.
 *
 * Decomposition rules (what FFTW's solvers do, re-planned for a GPU):
 *   - multi-D c2c: one batched 1-D transform per dimension, later ones in
 *     place on the output (dft/rank-geq2.c:42-52);
 *   - batch dimensions ride inside each kernel launch as up to three
 *     strided batch dims (dft/vrank-geq1.c:54-65);
 *   - 1-D length with a catalog kernel: one fused Stockham pass;
 *   - other smooth lengths n = n1*n2: column pass of n1 with the twiddle
 *     W_n^{k1*n2} fused into its store, then n2-point transforms written
 *     transposed (Cooley-Tukey, dft/ct.c:34-45, as a two-pass four-step);
 *   - lengths with large prime factors: Bluestein with a power-of-two
 *     convolution (dft/bluestein.c:82-128), fused into one kernel when the
 *     convolution fits one workgroup;
 *   - r2c: half-length complex transform + post-pass (rdft/ct-hc2c.c:
 *     197-210), then the leading dims as c2c on n/2+1 columns
 *     (rdft/rank-geq2-rdft2.c:40-53); c2r mirrors it (:55-69).
 */
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "engine.h"

/* ------------------------------------------------------------------ */
/* globals                                                             */

static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;
static const fk_kernel *g_reg = 0;
static int g_nreg = 0;
static int g_inited = 0;
static hipStream_t g_stream = 0;
static int g_async = 0;
static double g_timelimit = -1.0;
static __thread char g_err[512];

static void set_err(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char *fk_last_error(void) { return g_err; }
void fk_set_stream(hipStream_t s) { g_stream = s; }
hipStream_t fk_get_stream(void) { return g_stream; }
void fk_set_async(int on) { g_async = on; }
int fk_get_async(void) { return g_async; }
void fk_set_timelimit(double t) { g_timelimit = t; }

int fk_init(void)
{
    int i, ndev = 0;
    if (g_inited) return 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        set_err("no HIP device available");
        return -1;
    }
    g_reg = fk_registry(&g_nreg);
    for (i = 0; i < g_nreg; ++i) {
        if (g_reg[i].lds > 65536) {
            if (hipFuncSetAttribute(g_reg[i].fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)g_reg[i].lds) != hipSuccess) {
                set_err("hipFuncSetAttribute failed for L=%d", g_reg[i].L);
                (void)hipGetLastError();
            }
        }
    }
    /* device address of the allocation-free execute's static scratch on the
       current device (other devices take theirs on first use) */
    i = fk_static_scratch_setup();
    if (i) {
        set_err("static scratch set-up failed (%d)", i);
        (void)hipGetLastError();
    }
    g_inited = 1;
    return 0;
}

/* ------------------------------------------------------------------ */
/* device memory                                                       */

void *fk_malloc(size_t n)
{
    void *p = 0;
    if (n == 0) n = 1;
    if (hipMalloc(&p, n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return p;
}

void fk_free(void *p)
{
    if (p) (void)hipFree(p);
}

/* ------------------------------------------------------------------ */
/* cached device tables (freed by fk_cleanup)                          */

enum { T_STAGE, T_TW2, T_LINE, T_CHIRP, T_BLUEB };

typedef struct tbl {
    int type, prec, nrad;
    long long a, b;
    int rad[12];
    void *dev;
    size_t bytes;
    struct tbl *next;
} tbl;

static tbl *g_tbl = 0;

static tbl *tbl_find(int type, int prec, long long a, long long b, int nrad, const int *rad)
{
    tbl *t;
    for (t = g_tbl; t; t = t->next) {
        if (t->type != type || t->prec != prec || t->a != a || t->b != b || t->nrad != nrad)
            continue;
        if (nrad && memcmp(t->rad, rad, sizeof(int) * nrad)) continue;
        return t;
    }
    return 0;
}

static tbl *tbl_new(int type, int prec, long long a, long long b, int nrad, const int *rad,
                    size_t bytes)
{
    tbl *t = (tbl *)calloc(1, sizeof(tbl));
    t->type = type;
    t->prec = prec;
    t->a = a;
    t->b = b;
    t->nrad = nrad;
    if (nrad) memcpy(t->rad, rad, sizeof(int) * nrad);
    t->bytes = bytes;
    t->dev = fk_malloc(bytes);
    if (!t->dev) {
        free(t);
        return 0;
    }
    t->next = g_tbl;
    g_tbl = t;
    return t;
}

static size_t csz(int prec) { return prec == 0 ? 16 : 8; }
static size_t rsz(int prec) { return prec == 0 ? 8 : 4; }

static const void *stage_table(const fk_kernel *k, hipStream_t st)
{
    tbl *t;
    if (k->twsz == 0) return 0;
    t = tbl_find(T_STAGE, k->prec, k->L, 0, k->nrad, k->rad);
    if (t) return t->dev;
    t = tbl_new(T_STAGE, k->prec, k->L, 0, k->nrad, k->rad, csz(k->prec) * k->twsz);
    if (!t) return 0;
    if (fkm_gen_stage_tw(t->dev, k->prec, k->nrad, k->rad, k->twsz, st)) return 0;
    return t->dev;
}

static long long isqrt_ceil(long long n)
{
    long long b = (long long)sqrt((double)n);
    while (b * b < n) ++b;
    while (b > 1 && (b - 1) * (b - 1) >= n) --b;
    return b < 1 ? 1 : b;
}

/* FP64 two-level table of W_N^e: lo (b entries) then hi (N/b+1 entries) */
static const double *tw2_table(long long N, long long *b_out, hipStream_t st)
{
    /* split b ~ sqrt(N), a power of two when N is one (the kernels then
       index the two levels by mask and shift) */
    long long b = isqrt_ceil(N);
    if ((N & (N - 1)) == 0) {
        b = 1;
        while (b * b < N) b <<= 1;
    }
    tbl *t = tbl_find(T_TW2, 0, N, b, 0, 0);
    *b_out = b;
    if (t) return (const double *)t->dev;
    t = tbl_new(T_TW2, 0, N, b, 0, 0, 16 * (size_t)(b + N / b + 1));
    if (!t) return 0;
    if (fkm_gen_tw2((double *)t->dev, (double *)t->dev + 2 * b, N, b, st)) return 0;
    return (const double *)t->dev;
}

static const void *line_table(int prec, long long n, long long count, hipStream_t st)
{
    tbl *t = tbl_find(T_LINE, prec, n, count, 0, 0);
    if (t) return t->dev;
    t = tbl_new(T_LINE, prec, n, count, 0, 0, csz(prec) * (size_t)count);
    if (!t) return 0;
    if (fkm_gen_line_tw(t->dev, prec, n, count, st)) return 0;
    return t->dev;
}

static const void *chirp_table(int prec, long long n, hipStream_t st)
{
    tbl *t = tbl_find(T_CHIRP, prec, n, 0, 0, 0);
    if (t) return t->dev;
    t = tbl_new(T_CHIRP, prec, n, 0, 0, 0, csz(prec) * (size_t)n);
    if (!t) return 0;
    if (fkm_gen_chirp(t->dev, prec, n, st)) return 0;
    return t->dev;
}

static fk_plan *plan_internal(const fk_problem *pb, hipStream_t st);

/* B = DFT_M(conj chirp)/M, computed with an FP64 plan of the engine itself
   and rounded to the working precision */
static const void *blue_table(int prec, long long n, long long M, hipStream_t st)
{
    tbl *t = tbl_find(T_BLUEB, prec, n, M, 0, 0);
    fk_problem pb;
    fk_plan *sub;
    void *b64, *B64;
    int rc;
    if (t) return t->dev;
    b64 = fk_malloc(16 * (size_t)M);
    B64 = fk_malloc(16 * (size_t)M);
    if (!b64 || !B64) goto fail;
    memset(&pb, 0, sizeof(pb));
    pb.kind = FK_C2C;
    pb.prec = 0;
    pb.sign = -1;
    pb.flags = FK_ESTIMATE;
    pb.rank = 1;
    pb.dims[0].n = M;
    pb.dims[0].is = pb.dims[0].os = 1;
    pb.in = b64;
    pb.out = B64;
    sub = plan_internal(&pb, st);
    if (!sub) goto fail;
    sub->stream = st;
    rc = fkm_gen_blue_b(b64, n, M, st);
    if (!rc) rc = fk_execute(sub, b64, B64);
    if (rc) {
        fk_plan_destroy(sub);
        goto fail;
    }
    t = tbl_new(T_BLUEB, prec, n, M, 0, 0, csz(prec) * (size_t)M);
    if (t) rc = fkm_scale_round(t->dev, prec, B64, M, 1.0 / (double)M, st);
    (void)hipStreamSynchronize(st);
    fk_plan_destroy(sub);
    fk_free(b64);
    fk_free(B64);
    return t && !rc ? t->dev : 0;
fail:
    fk_free(b64);
    fk_free(B64);
    return 0;
}

/* ------------------------------------------------------------------ */
/* kernel catalog queries                                              */

static const fk_kernel *find_kernel(int prec, long long L, int mode, int io)
{
    int i;
    if (L > (1 << 20)) return 0;
    for (i = 0; i < g_nreg; ++i) {
        const fk_kernel *k = g_reg + i;
        if (k->prec == prec && k->L == L && k->mode == mode && k->io == io) return k;
    }
    return 0;
}

static int has_single(int prec, long long n) { return find_kernel(prec, n, FK_ROW, FK_IO_PLAIN) != 0; }

/* the catalog variant of (prec, L, mode, io) whose stages are all radix 16
   (the matrix-core substitution needs them), or NULL */
static const fk_kernel *find_kernel_r16(int prec, long long L, int mode, int io)
{
    int i, j;
    for (i = 0; i < g_nreg; ++i) {
        const fk_kernel *k = g_reg + i;
        int all16 = k->nrad > 0;
        if (k->prec != prec || k->L != L || k->mode != mode || k->io != io) continue;
        for (j = 0; j < k->nrad; ++j) all16 &= k->rad[j] == 16;
        if (all16) return k;
    }
    return 0;
}

/* n decomposable into catalog passes without Bluestein */
static int is_pow16(long long n)
{
    while (n > 1 && n % 16 == 0) n /= 16;
    return n == 1;
}

static int smooth_plannable(int prec, long long n, int depth)
{
    long long n1;
    if (n == 1 || has_single(prec, n)) return 1;
    if (depth >= 3) return 0;
    for (n1 = 2; n1 <= 16384 && n1 < n; ++n1) {
        if (n % n1) continue;
        if (!find_kernel(prec, n1, FK_COL, FK_IO_PLAIN)) continue;
        if (smooth_plannable(prec, n / n1, depth + 1)) return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* plan building                                                       */

/* Planner decision kinds, ordered by the patience level that explores them
   (MEASURE explores kernel geometries and thread mappings, PATIENT also
   alternative two-pass splits). */
enum { DEC_VARIANT = 1, DEC_MODE = 1, DEC_SPLIT = 2, DEC_CHUNK = 2 };
#define FK_MAX_DECIDE 32

typedef struct {
    fk_plan *P;
    hipStream_t st;
    int err;
    const int *choice;          /* decisions to replay (NULL: all defaults) */
    int nchoice;
    int ci;                     /* decisions taken so far */
    int r16;                    /* restrict to radix-16 passes (FK_MFMA) */
    int bin, bout;              /* buffers of the r2c / c2r problem's arrays */
    int nalt[FK_MAX_DECIDE];    /* alternatives at each decision */
    int dkind[FK_MAX_DECIDE];
} ctx;

static int decide(ctx *c, int nalt, int kind)
{
    int i = c->ci++, v = 0;
    if (i < FK_MAX_DECIDE && c->choice && i < c->nchoice) v = c->choice[i];
    if (v >= nalt) v = nalt - 1;
    if (v < 0) v = 0;
    if (i < FK_MAX_DECIDE) {
        c->nalt[i] = nalt;
        c->dkind[i] = kind;
        c->P->choice[i] = v;
    }
    return v;
}

/* all catalog variants of (prec, L, mode, io), default first */
static int find_variants(int prec, long long L, int mode, int io, const fk_kernel **out, int max)
{
    int i, n = 0, v;
    if (L > (1 << 20)) return 0;
    for (v = 0; v < max; ++v) {
        const fk_kernel *hit = 0;
        for (i = 0; i < g_nreg; ++i) {
            const fk_kernel *k = g_reg + i;
            if (k->prec == prec && k->L == L && k->mode == mode && k->io == io && k->variant == v) {
                hit = k;
                break;
            }
        }
        if (!hit) break;
        out[n++] = hit;
    }
    return n;
}

/* Catalog entry forced by fk_force_kernel (NULL: none).  Wherever the
   planner picks a kernel of the same precision, length, mapping and IO kind
   it takes this one instead, so that every catalog entry, including the
   variants only the autotuner would pick, can be run and checked. */
static const fk_kernel *g_force = 0;

static int forced(int prec, long long L, int mode, int io)
{
    const fk_kernel *f = g_force;
    return f && f->prec == prec && f->L == L && f->mode == mode && f->io == io;
}

static const fk_kernel *pick_variant(ctx *c, int prec, long long L, int mode, int io)
{
    const fk_kernel *vs[16];
    int n = find_variants(prec, L, mode, io, vs, 16), v;
    if (!n) return 0;
    v = n > 1 ? decide(c, n, DEC_VARIANT) : 0;
    return forced(prec, L, mode, io) ? g_force : vs[v];
}

int fk_catalog_size(void)
{
    if (!g_reg) g_reg = fk_registry(&g_nreg);
    return g_nreg;
}

const fk_kernel *fk_catalog_entry(int i)
{
    if (!g_reg) g_reg = fk_registry(&g_nreg);
    return i >= 0 && i < g_nreg ? g_reg + i : 0;
}

int fk_force_kernel(int i)
{
    const fk_kernel *k = i < 0 ? 0 : fk_catalog_entry(i);
    if (i >= 0 && !k) return -1;
    pthread_mutex_lock(&g_lock);
    g_force = k;
    pthread_mutex_unlock(&g_lock);
    return 0;
}

/* Workgroups of kernel fn (block threads, lds bytes) that the current
   device holds at once over all its CUs: the grid of the kernels whose
   workgroups run several tiles.  0 when the runtime cannot say (the caller
   keeps one workgroup per tile). */
static unsigned resident_wgs(const void *fn, unsigned block, unsigned lds)
{
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, (int)block, lds) != hipSuccess ||
        ncu <= 0 || per <= 0) {
        (void)hipGetLastError();
        return 0;
    }
    return (unsigned)ncu * (unsigned)per;
}

static fk_pass *new_pass(ctx *c)
{
    fk_plan *P = c->P;
    if (P->npass == P->cpass) {
        P->cpass = P->cpass ? 2 * P->cpass : 8;
        P->pass = (fk_pass *)realloc(P->pass, sizeof(fk_pass) * P->cpass);
    }
    memset(P->pass + P->npass, 0, sizeof(fk_pass));
    return P->pass + P->npass++;
}

static void need_scratch(ctx *c, int slot, size_t bytes)
{
    fk_plan *P = c->P;
    if (slot >= FK_MAX_SCRATCH) {
        c->err = 1;
        set_err("too many scratch buffers");
        return;
    }
    if (P->scratch_bytes[slot] < bytes) P->scratch_bytes[slot] = bytes;
    if (P->nscratch < slot + 1) P->nscratch = slot + 1;
}

/* a pass reads or writes buffer `buf` in units of `a` bytes: the array bound
   to it at execute must keep that alignment (staging scratch always does) */
static void need_align(ctx *c, int buf, unsigned a)
{
    if (buf == FK_BUF_IN && c->P->align_in < a) c->P->align_in = a;
    if (buf == FK_BUF_OUT && c->P->align_out < a) c->P->align_out = a;
}

static long long llabs_(long long x) { return x < 0 ? -x : x; }

/* Drop unit dims; stable. */
static int squeeze(int nb, fk_dim *b)
{
    int i, m = 0;
    for (i = 0; i < nb; ++i)
        if (b[i].n != 1) b[m++] = b[i];
    return m;
}

static void sort_by_is(int nb, fk_dim *b)
{
    int i, j;
    for (i = 1; i < nb; ++i)
        for (j = i; j > 0 && llabs_(b[j].is) < llabs_(b[j - 1].is); --j) {
            fk_dim t = b[j];
            b[j] = b[j - 1];
            b[j - 1] = t;
        }
}

/* Merge batch dims that describe one contiguous run on both sides. */
static int merge_dims(int nb, fk_dim *b)
{
    int i, changed = 1;
    while (changed) {
        changed = 0;
        for (i = 0; i + 1 < nb && !changed; ++i) {
            int j;
            for (j = 0; j < nb; ++j) {
                if (j == i) continue;
                /* dim j directly outside dim i? */
                if (b[j].is == b[i].is * b[i].n && b[j].os == b[i].os * b[i].n) {
                    b[i].n *= b[j].n;
                    memmove(b + j, b + j + 1, sizeof(fk_dim) * (nb - j - 1));
                    --nb;
                    changed = 1;
                    break;
                }
            }
        }
    }
    return nb;
}

/* Emit one pass per combination of batch dims beyond the three a kernel
   addresses itself. */
typedef void (*emit_fn)(ctx *c, fk_pass *proto, int nb, const fk_dim *b, long long ioff,
                        long long ooff);

static void fill_batch3(long long *nbv, long long *ids, long long *ods, long long *nl, int nb,
                        const fk_dim *b)
{
    int i;
    long long t = 1;
    for (i = 0; i < 3; ++i) {
        if (i < nb) {
            nbv[i] = b[i].n;
            ids[i] = b[i].is;
            ods[i] = b[i].os;
        } else {
            nbv[i] = 1;
            ids[i] = ods[i] = 0;
        }
        t *= nbv[i];
    }
    *nl = t;
}

static void emit_loops(ctx *c, const fk_pass *proto, int nb, const fk_dim *b0, size_t isz,
                       size_t osz)
{
    long long idx[FK_MAXRANK * 2] = {0};
    int nloop = nb > 3 ? nb - 3 : 0, i, j;
    fk_dim b[FK_MAXRANK * 2];
    memcpy(b, b0, sizeof(fk_dim) * nb);
    /* beyond three batch dims, keep the two largest after b[0] (whose place
       the thread mapping may depend on) in the kernel and loop over the
       smallest: fewer, larger launches */
    for (i = 2; nb > 3 && i < nb; ++i)
        for (j = i; j > 1 && b[j].n > b[j - 1].n; --j) {
            fk_dim t = b[j];
            b[j] = b[j - 1];
            b[j - 1] = t;
        }
    for (;;) {
        long long ioff = proto->ioff, ooff = proto->ooff;
        long long i2off = proto->i2off, o2off = proto->o2off;
        fk_pass *p;
        for (i = 0; i < nloop; ++i) {
            ioff += idx[i] * b[3 + i].is * (long long)isz;
            ooff += idx[i] * b[3 + i].os * (long long)osz;
            i2off += idx[i] * b[3 + i].is * (long long)isz;
            o2off += idx[i] * b[3 + i].os * (long long)osz;
        }
        p = new_pass(c);
        *p = *proto;
        p->ioff = ioff;
        p->ooff = ooff;
        p->i2off = i2off;
        p->o2off = o2off;
        if (p->type == FK_PASS_STOCKHAM) {
            fill_batch3(p->a.nb, p->a.ids, p->a.ods, &p->a.nlines, nb < 3 ? nb : 3, b);
            p->grid = (unsigned)((p->a.nlines + p->k->nl - 1) / p->k->nl);
            if (p->k->loop) {
                unsigned cap = resident_wgs(p->k->fn, p->block, p->lds);
                if (cap && cap < p->grid) p->grid = cap;
            }
            if (p->a.nlines == 0) --c->P->npass;
        } else {
            fill_batch3(p->e.nb, p->e.ids, p->e.ods, &p->e.nlines, nb < 3 ? nb : 3, b);
            if (p->e.nlines == 0) --c->P->npass;
        }
        for (i = 0; i < nloop; ++i) {
            if (++idx[i] < b[3 + i].n) break;
            idx[i] = 0;
        }
        if (i == nloop) break;
    }
}

static double pass_bytes_c(int prec, double n_elems) { return 2.0 * n_elems * csz(prec); }

static long long batch_count(int nb, const fk_dim *b)
{
    long long t = 1;
    int i;
    for (i = 0; i < nb; ++i) t *= b[i].n;
    return t;
}

static void emit_stockham(ctx *c, const fk_kernel *k, long long is, long long os, int nb,
                          const fk_dim *b, int ibuf, long long ioff, int obuf, long long ooff,
                          unsigned flags, size_t isz, size_t osz)
{
    fk_pass proto;
    if (k->loop && (flags & FK_TW_OUT)) {
        /* the pipelined kernels have no four-step output twiddle: such
           passes run the one-tile-per-workgroup kernel of the same
           geometry */
        const fk_kernel *vs[16];
        int n = find_variants(k->prec, k->L, k->mode, k->io, vs, 16), i, j, same;
        for (i = 0; i < n; ++i) {
            if (vs[i]->loop || vs[i]->tpl != k->tpl || vs[i]->nl != k->nl || vs[i]->nrad != k->nrad)
                continue;
            for (same = 1, j = 0; j < k->nrad; ++j) same &= vs[i]->rad[j] == k->rad[j];
            if (same) {
                k = vs[i];
                break;
            }
        }
        if (k->loop) {
            c->err = 1;
            set_err("no twiddled kernel for L=%d", k->L);
            return;
        }
    }
    memset(&proto, 0, sizeof(proto));
    proto.type = FK_PASS_STOCKHAM;
    proto.k = k;
    proto.ibuf = ibuf;
    proto.obuf = obuf;
    proto.ioff = ioff;
    proto.ooff = ooff;
    proto.block = (unsigned)(k->tpl * k->nl);
    proto.lds = k->lds;
    proto.a.is = is;
    proto.a.os = os;
    proto.a.flags = flags;
    proto.a.tw = stage_table(k, c->st);
    if (k->twsz && !proto.a.tw) {
        c->err = 1;
        set_err("twiddle table allocation failed");
        return;
    }
    proto.a.xcd_remap = (k->mode == FK_COL || k->io == FK_IO_TOUT) ? 1u : 0u;
    proto.bytes = pass_bytes_c(k->prec, (double)k->L * (double)batch_count(nb, b));
    emit_loops(c, &proto, nb, b, isz, osz);
}

static int plan_c2c_1d(ctx *c, long long n, long long is, long long os, int nb,
                       const fk_dim *bin, int ib, long long io, int ob, long long oo, int swi,
                       int swo, int slot);

static int plan_bluestein(ctx *c, long long n, long long is, long long os, int nb,
                          const fk_dim *b, int ib, long long io, int ob, long long oo, int swi,
                          int swo, int slot)
{
    int prec = c->P->prec;
    long long M = 1;
    const fk_kernel *kb;
    const void *chirp, *B;
    fk_dim bb[FK_MAXRANK * 2];
    size_t cs = csz(prec);
    while (M < 2 * n - 1) M <<= 1;
    chirp = chirp_table(prec, n, c->st);
    B = blue_table(prec, n, M, c->st);
    if (!chirp || !B) {
        c->err = 1;
        set_err("Bluestein tables failed (n=%lld)", n);
        return -1;
    }
    memcpy(bb, b, sizeof(fk_dim) * nb);
    kb = forced(prec, M, FK_ROW, FK_IO_BLUE) ? g_force : find_kernel(prec, M, FK_ROW, FK_IO_BLUE);
    if (kb) {
        size_t before = (size_t)c->P->npass;
        unsigned fl = (swi ? FK_SWAP_IN : 0) | (swo ? FK_SWAP_OUT : 0);
        emit_stockham(c, kb, is, os, nb, bb, ib, io, ob, oo, fl, cs, cs);
        for (; before < (size_t)c->P->npass; ++before) {
            fk_pass *p = c->P->pass + before;
            p->a.aux = chirp;
            p->a.aux2 = B;
            p->a.nvalid = n;
            p->bytes = pass_bytes_c(prec, (double)n * (double)batch_count(nb, b));
        }
        return c->err ? -1 : 0;
    } else {
        /* multi-pass: prep -> DFT_M -> pointwise -> DFT_M -> post */
        long long H = batch_count(nb, b), s;
        fk_dim sb[FK_MAXRANK * 2];
        fk_pass proto;
        int i, s1 = slot, s2 = slot + 1;
        need_scratch(c, s1, cs * (size_t)M * (size_t)H);
        need_scratch(c, s2, cs * (size_t)M * (size_t)H);
        /* scratch batch layout: contiguous [h][M] in the order of b */
        s = M;
        for (i = 0; i < nb; ++i) {
            sb[i].n = b[i].n;
            sb[i].is = b[i].is;
            sb[i].os = s;
            s *= b[i].n;
        }
        memset(&proto, 0, sizeof(proto));
        proto.type = FK_PASS_EW;
        proto.e.op = FK_EW_BS_PREP;
        proto.e.n = M;
        proto.e.is = is;
        proto.e.os = 1;
        proto.e.aux = chirp;
        proto.e.p0 = n;
        proto.e.p1 = M;
        proto.e.flags = swi ? FK_SWAP_IN : 0;
        proto.ibuf = ib;
        proto.ioff = io;
        proto.obuf = FK_BUF_S0 + s1;
        proto.bytes = (double)cs * (double)(n + M) * (double)H;
        emit_loops(c, &proto, nb, sb, cs, cs);
        for (i = 0; i < nb; ++i) sb[i].is = sb[i].os;
        if (plan_c2c_1d(c, M, 1, 1, nb, sb, FK_BUF_S0 + s1, 0, FK_BUF_S0 + s2, 0, 0, 0, slot + 2))
            return -1;
        memset(&proto, 0, sizeof(proto));
        proto.type = FK_PASS_EW;
        proto.e.op = FK_EW_BS_MUL;
        proto.e.n = M;
        proto.e.is = proto.e.os = 1;
        proto.e.aux2 = B;
        proto.ibuf = proto.obuf = FK_BUF_S0 + s2;
        proto.bytes = 2.0 * (double)cs * (double)M * (double)H;
        emit_loops(c, &proto, nb, sb, cs, cs);
        if (plan_c2c_1d(c, M, 1, 1, nb, sb, FK_BUF_S0 + s2, 0, FK_BUF_S0 + s1, 0, 0, 0, slot + 2))
            return -1;
        for (i = 0; i < nb; ++i) {
            sb[i].is = sb[i].os;
            sb[i].os = b[i].os;
        }
        memset(&proto, 0, sizeof(proto));
        proto.type = FK_PASS_EW;
        proto.e.op = FK_EW_BS_POST;
        proto.e.n = n;
        proto.e.is = 1;
        proto.e.os = os;
        proto.e.aux = chirp;
        proto.e.p0 = n;
        proto.e.flags = swo ? FK_SWAP_OUT : 0;
        proto.ibuf = FK_BUF_S0 + s1;
        proto.obuf = ob;
        proto.ooff = oo;
        proto.bytes = 2.0 * (double)cs * (double)n * (double)H;
        emit_loops(c, &proto, nb, sb, cs, cs);
        return c->err ? -1 : 0;
    }
}

/* pick n1 (column pass length) for a two-pass split of n */
/* Relative cost of an HBM pass whose contiguous runs are `seg` bytes
   (one 128-byte line and up stream well; shorter runs waste bursts). */
static double seg_pen(double seg)
{
    return seg >= 256 ? 0.0 : seg >= 128 ? 0.15 : seg >= 64 ? 0.5 : seg >= 32 ? 1.5 : 3.0;
}

/* Candidate column lengths n1 for a split of n, best first.  Score = modelled
   HBM passes: the column pass costs 1 + seg_pen(lines per tile x element),
   a single-pass second factor 1 + seg_pen of its transposed store, a
   further split about 2.3; balance breaks ties.  Returns the number of
   candidates (<= max). */
static int split_candidates(int prec, long long n, long long *out, int max)
{
    long long n1;
    double sc[64];
    long long cand[64];
    int nc = 0, i, j;
    for (n1 = 2; n1 <= 16384 && n1 < n; ++n1) {
        long long n2;
        double score;
        const fk_kernel *kc, *kt;
        if (n % n1) continue;
        n2 = n / n1;
        kc = find_kernel(prec, n1, FK_COL, FK_IO_PLAIN);
        if (!kc) continue;
        if (!smooth_plannable(prec, n2, 1)) continue;
        score = 1.0 + seg_pen((double)kc->nl * csz(prec));
        if (has_single(prec, n2)) {
            kt = find_kernel(prec, n2, FK_ROW, FK_IO_TOUT);
            score += 1.0 + (kt ? seg_pen((double)kt->nl * csz(prec)) : 3.0);
        } else {
            score += 2.3;
        }
        score += 0.05 * fabs(log((double)n1) - log((double)n2));
        score -= 1e-9 * (double)n2; /* ties: larger second (row) factor */
        if (nc < 64) {
            sc[nc] = score;
            cand[nc++] = n1;
        }
    }
    for (i = 1; i < nc; ++i)
        for (j = i; j > 0 && sc[j] < sc[j - 1]; --j) {
            double t = sc[j];
            long long u = cand[j];
            sc[j] = sc[j - 1];
            cand[j] = cand[j - 1];
            sc[j - 1] = t;
            cand[j - 1] = u;
        }
    if (nc > max) nc = max;
    for (i = 0; i < nc; ++i) out[i] = cand[i];
    return nc;
}

/*
 * Batched 1-D complex DFT of length n: element j of batch line h at
 * h-offset + j*is (input) / j*os (output).  Strides in complex elements.
 */
static int plan_c2c_1d(ctx *c, long long n, long long is, long long os, int nb_,
                       const fk_dim *bin, int ib, long long io, int ob, long long oo, int swi,
                       int swo, int slot)
{
    int prec = c->P->prec, nb, i;
    size_t cs = csz(prec);
    fk_dim b[FK_MAXRANK * 2];
    const fk_kernel *kr, *kc, *kt;
    unsigned fl = (swi ? FK_SWAP_IN : 0) | (swo ? FK_SWAP_OUT : 0);
    memcpy(b, bin, sizeof(fk_dim) * nb_);
    nb = squeeze(nb_, b);
    nb = merge_dims(nb, b);
    sort_by_is(nb, b);
    if (n == 1) {
        /* length-1 DFT is a copy (with the re/im swaps cancelling) */
        fk_pass proto;
        if (ib == ob && io == oo && is == os) {
            int same = 1;
            for (i = 0; i < nb; ++i) same &= b[i].is == b[i].os;
            if (same) return 0;
        }
        memset(&proto, 0, sizeof(proto));
        proto.type = FK_PASS_EW;
        proto.e.op = FK_EW_COPY;
        proto.e.n = 1;
        proto.e.is = is;
        proto.e.os = os;
        proto.e.flags = fl;
        proto.ibuf = ib;
        proto.ioff = io;
        proto.obuf = ob;
        proto.ooff = oo;
        proto.bytes = 2.0 * cs * (double)batch_count(nb, b);
        emit_loops(c, &proto, nb, b, cs, cs);
        return 0;
    }
    kr = find_kernel(prec, n, FK_ROW, FK_IO_PLAIN);
    if (c->r16 && n != 16 && n != 256 && n != 4096) kr = 0;
    if (kr) {
        /* candidate thread mappings, the heuristic's choice first:
           0 ROW, 1 ROW with transposed store, 2 COL */
        int cand[3], ncand = 0, dT = -1, dC = -1, m;
        long long best = 1LL << 62;
        const fk_kernel *k;
        kc = find_kernel(prec, n, FK_COL, FK_IO_PLAIN);
        kt = find_kernel(prec, n, FK_ROW, FK_IO_TOUT);
        if (kt && llabs_(os) > 1)
            for (i = 0; i < nb; ++i)
                if (llabs_(b[i].os) == 1) { dT = i; break; }
        for (i = 0; i < nb; ++i) {
            long long s = llabs_(b[i].is) + llabs_(b[i].os);
            if (s < best && b[i].n >= 2) { best = s; dC = i; }
        }
        if (!kc || best > 4) dC = -1;
        if (llabs_(is) == 1) {
            if (dT >= 0) cand[ncand++] = 1;
            cand[ncand++] = 0;
            if (dC >= 0) cand[ncand++] = 2;
        } else {
            if (dC >= 0) cand[ncand++] = 2;
            cand[ncand++] = 0;
            if (dT >= 0) cand[ncand++] = 1;
        }
        m = cand[ncand > 1 ? decide(c, ncand, DEC_MODE) : 0];
        if (m == 1 || m == 2) {
            int d = m == 1 ? dT : dC;
            fk_dim t = b[d];
            memmove(b + 1, b, sizeof(fk_dim) * d);
            b[0] = t;
        }
        if (m == 1) k = pick_variant(c, prec, n, FK_ROW, FK_IO_TOUT);
        else if (m == 2) k = pick_variant(c, prec, n, FK_COL, FK_IO_PLAIN);
        else k = pick_variant(c, prec, n, FK_ROW, FK_IO_PLAIN);
        if (c->r16) {
            /* matrix-core plans: the all-radix-16 variant of the mapping */
            const fk_kernel *k16 = find_kernel_r16(prec, n, k->mode, k->io);
            if (k16) k = k16;
        }
        emit_stockham(c, k, is, os, nb, b, ib, io, ob, oo, fl, cs, cs);
        return c->err ? -1 : 0;
    }
    if (smooth_plannable(prec, n, 0)) {
        long long cands[4], n1, n2, H, s;
        int ncand = split_candidates(prec, n, cands, 4);
        if (c->r16 && is_pow16(n)) {
            /* matrix-core variant: radix-16 passes only (16, 256, 4096) */
            ncand = 0;
            /* 256 first: its column passes hold 16 lines per tile (128-byte
               runs), where a 4096-point column pass holds 2 */
            if (n % 256 == 0 && is_pow16(n / 256)) cands[ncand++] = 256;
            if (n % 4096 == 0 && is_pow16(n / 4096)) cands[ncand++] = 4096;
            if (is_pow16(n / 16)) cands[ncand++] = 16;
        }
        fk_dim ab[FK_MAXRANK * 2 + 1], bb[FK_MAXRANK * 2 + 1];
        fk_pass proto;
        const fk_kernel *kA;
        size_t before;
        if (!ncand) goto blue;
        n1 = cands[ncand > 1 ? decide(c, ncand, DEC_SPLIT) : 0];
        n2 = n / n1;
        kA = pick_variant(c, prec, n1, FK_COL, FK_IO_PLAIN);
        H = batch_count(nb, b);
        need_scratch(c, slot, cs * (size_t)n * (size_t)H);
        /* pass A: lines n2_idx (tile dim), elements n1_idx at stride is*n2 */
        ab[0].n = n2;
        ab[0].is = is;
        ab[0].os = 1;
        s = n;
        for (i = 0; i < nb; ++i) {
            ab[i + 1].n = b[i].n;
            ab[i + 1].is = b[i].is;
            ab[i + 1].os = s;
            s *= b[i].n;
        }
        memset(&proto, 0, sizeof(proto));
        before = (size_t)c->P->npass;
        emit_stockham(c, kA, is * n2, n2, nb + 1, ab, ib, io, FK_BUF_S0 + slot, 0,
                      (swi ? FK_SWAP_IN : 0) | FK_TW_OUT, cs, cs);
        if (c->err) return -1;
        {
            long long tb;
            const double *t = tw2_table(n, &tb, c->st);
            if (!t) { c->err = 1; return -1; }
            for (; before < (size_t)c->P->npass; ++before) {
                fk_pass *p = c->P->pass + before;
                p->a.tlo = t;
                p->a.thi = t + 2 * tb;
                p->a.tN = n;
                p->a.tb = tb;
            }
        }
        /* pass B: n2-point transforms of rows k1, written to k1 + n1*k2 */
        bb[0].n = n1;
        bb[0].is = n2;
        bb[0].os = os;
        for (i = 0; i < nb; ++i) {
            bb[i + 1].n = b[i].n;
            bb[i + 1].is = ab[i + 1].os;
            bb[i + 1].os = b[i].os;
        }
        return plan_c2c_1d(c, n2, 1, n1 * os, nb + 1, bb, FK_BUF_S0 + slot, 0, ob, oo, 0, swo,
                           slot + 1);
    }
blue:
    return plan_bluestein(c, n, is, os, nb, b, ib, io, ob, oo, swi, swo, slot);
}

/* ------------------------------------------------------------------ */
/* problem-level planning                                              */

static void desc_add(fk_plan *P, const char *fmt, ...)
{
    size_t l = strlen(P->desc);
    va_list ap;
    if (l + 8 >= sizeof(P->desc)) return;
    va_start(ap, fmt);
    vsnprintf(P->desc + l, sizeof(P->desc) - l, fmt, ap);
    va_end(ap);
}

/* c2c over all transform dims, from (ib) to (ob), then in place on ob */
static int plan_c2c_dims(ctx *c, int rank, const fk_dim *dims, int vrank, const fk_dim *vdims,
                         int ib, int ob, int sw)
{
    int d, cur = ib, i;
    for (d = rank - 1; d >= 0; --d) {
        fk_dim b[FK_MAXRANK * 2];
        int nb = 0;
        long long is = cur == ib ? dims[d].is : dims[d].os;
        for (i = 0; i < rank; ++i) {
            if (i == d) continue;
            b[nb].n = dims[i].n;
            b[nb].is = cur == ib ? dims[i].is : dims[i].os;
            b[nb].os = dims[i].os;
            ++nb;
        }
        for (i = 0; i < vrank; ++i) {
            b[nb].n = vdims[i].n;
            b[nb].is = cur == ib ? vdims[i].is : vdims[i].os;
            b[nb].os = vdims[i].os;
            ++nb;
        }
        if (plan_c2c_1d(c, dims[d].n, is, dims[d].os, nb, b, cur, 0, ob, 0, sw, sw, 0))
            return -1;
        cur = ob;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* cache blocking of multi-dimensional plans                           */
/*
 * MI355X's 256 MB memory-side cache (Infinity Cache) keeps a working set of
 * that order between two passes.  A multi-dimensional plan can therefore
 * run its passes chunk by chunk, so that the intermediate of a chunk is
 * re-read from that cache instead of HBM:
 *   - batched problems (rank >= 2, howmany > 1): every pass, per chunk of
 *     transforms;
 *   - unbatched rank >= 3 problems: the passes of dims 1..rank-1 per slab
 *     of dim-0 planes, then the dim-0 pass over the whole array.
 * The chunk size is a planner decision timed under FFTW_PATIENT and
 * FFTW_EXHAUSTIVE (on the 512^3 FP64 and 2048^2 x 64 FP32 shapes the
 * whole-problem plan measured faster: 2.82 vs 3.12 ms, 1.19 vs 1.40 ms);
 * ESTIMATE and MEASURE plan the whole problem at once unless
 * FFTW_AMD_BLOCKED asks for blocking.
 */
static const long long chunk_targets[] = {64LL << 20, 128LL << 20, 192LL << 20};

/* Offsets the arrays' byte offsets of passes [from, npass) (the chunk's
   position in the caller's arrays). */
static void shift_passes(fk_plan *P, int from, long long in_off, long long out_off)
{
    int i;
    for (i = from; i < P->npass; ++i) {
        fk_pass *p = P->pass + i;
        if (p->ibuf == FK_BUF_IN) p->ioff += in_off;
        else if (p->ibuf == FK_BUF_OUT) p->ioff += out_off;
        if (p->obuf == FK_BUF_IN) p->ooff += in_off;
        else if (p->obuf == FK_BUF_OUT) p->ooff += out_off;
    }
}

/* Units (transforms or planes) per chunk: all of them, or as many as fit
   a cache target; a decision when there is a choice. */
static long long chunk_units(ctx *c, long long nunits, double unit_bytes)
{
    long long cand[4];
    int nc = 0, t;
    cand[nc++] = nunits;
    for (t = 0; t < (int)(sizeof(chunk_targets) / sizeof(*chunk_targets)); ++t) {
        long long k = (long long)((double)chunk_targets[t] / unit_bytes);
        if (k < 1) k = 1;
        if (k < nunits && k != cand[nc - 1]) cand[nc++] = k;
    }
    if (nc == 1) return cand[0];
    if ((c->P->flags & FK_BLOCKED) && !(c->choice && c->ci < c->nchoice)) {
        /* FFTW_AMD_BLOCKED: the first cache-sized chunk is the default */
        int i = c->ci++;
        if (i < FK_MAX_DECIDE) {
            c->nalt[i] = nc;
            c->dkind[i] = DEC_CHUNK;
            c->P->choice[i] = 1;
        }
        return cand[1];
    }
    return cand[decide(c, nc, DEC_CHUNK)];
}

typedef int (*chunk_fn)(ctx *c, const fk_problem *q, int sw);

/* Plans q in chunks of `k` indices of vdim v (balanced chunks); every
   chunk replays the decisions of the first. */
static int plan_vchunks(ctx *c, const fk_problem *pb, int v, long long k, chunk_fn f, int sw,
                        size_t isz, size_t osz)
{
    long long n = pb->vdims[v].n, nch = (n + k - 1) / k, h0 = 0, j;
    int ci0 = c->ci, ci1 = c->ci;
    for (j = 0; j < nch; ++j) {
        fk_problem q = *pb;
        long long cnt = n / nch + (j < n % nch);
        int before = c->P->npass;
        q.vdims[v].n = cnt;
        c->ci = ci0;
        if (f(c, &q, sw)) return -1;
        if (c->ci > ci1) ci1 = c->ci;
        shift_passes(c->P, before, h0 * pb->vdims[v].is * (long long)isz,
                     h0 * pb->vdims[v].os * (long long)osz);
        h0 += cnt;
    }
    c->ci = ci1;
    return 0;
}

static int chunk_c2c(ctx *c, const fk_problem *q, int sw)
{
    return plan_c2c_dims(c, q->rank, q->dims, q->vrank, q->vdims, FK_BUF_IN, FK_BUF_OUT, sw);
}

static int plan_r2c(ctx *c, const fk_problem *pb);
static int plan_c2r(ctx *c, const fk_problem *pb);
static int chunk_r2c(ctx *c, const fk_problem *q, int sw) { (void)sw; return plan_r2c(c, q); }
static int chunk_c2r(ctx *c, const fk_problem *q, int sw) { (void)sw; return plan_c2r(c, q); }

/* vdim to chunk over: the one with the most indices */
static int chunk_vdim(const fk_problem *pb)
{
    int i, v = -1;
    for (i = 0; i < pb->vrank; ++i)
        if (pb->vdims[i].n > 1 && (v < 0 || pb->vdims[i].n > pb->vdims[v].n)) v = i;
    return v;
}

/* Cache-blocked c2c (rank >= 2); returns 1 if it planned the problem, 0 if
   the whole-problem plan is the choice, -1 on error. */
static int plan_c2c_blocked(ctx *c, const fk_problem *pb, int sw)
{
    fk_plan *P = c->P;
    size_t cs = csz(P->prec);
    long long N = 1, H = 1, k;
    int i, v;
    for (i = 0; i < pb->rank; ++i) N *= pb->dims[i].n;
    for (i = 0; i < pb->vrank; ++i) H *= pb->vdims[i].n;
    if (pb->rank < 2) return 0;
    v = chunk_vdim(pb);
    if (v >= 0) {
        k = chunk_units(c, pb->vdims[v].n, (double)N * cs * (double)(H / pb->vdims[v].n));
        if (k >= pb->vdims[v].n) return 0;
        desc_add(P, "(blocked %lld of %lld)", k, pb->vdims[v].n);
        return plan_vchunks(c, pb, v, k, chunk_c2c, sw, cs, cs) ? -1 : 1;
    }
    if (pb->rank < 3) return 0;
    {
        long long n0 = pb->dims[0].n, nch, h0 = 0, j;
        int ci0, ci1, nb = 0;
        fk_dim b[FK_MAXRANK * 2];
        k = chunk_units(c, n0, (double)(N / n0) * cs);
        if (k >= n0) return 0;
        desc_add(P, "(blocked slabs %lld of %lld)", k, n0);
        nch = (n0 + k - 1) / k;
        ci0 = ci1 = c->ci;
        for (j = 0; j < nch; ++j) {
            fk_dim vd[FK_MAXRANK + 1];
            long long cnt = n0 / nch + (j < n0 % nch);
            int before = P->npass;
            memcpy(vd, pb->vdims, sizeof(fk_dim) * pb->vrank);
            vd[pb->vrank].n = cnt;
            vd[pb->vrank].is = pb->dims[0].is;
            vd[pb->vrank].os = pb->dims[0].os;
            c->ci = ci0;
            if (plan_c2c_dims(c, pb->rank - 1, pb->dims + 1, pb->vrank + 1, vd, FK_BUF_IN,
                              FK_BUF_OUT, sw))
                return -1;
            if (c->ci > ci1) ci1 = c->ci;
            shift_passes(P, before, h0 * pb->dims[0].is * (long long)cs,
                         h0 * pb->dims[0].os * (long long)cs);
            h0 += cnt;
        }
        c->ci = ci1;
        /* dim 0 over the whole array, in place on the output */
        for (i = 1; i < pb->rank; ++i) {
            b[nb].n = pb->dims[i].n;
            b[nb].is = b[nb].os = pb->dims[i].os;
            ++nb;
        }
        for (i = 0; i < pb->vrank; ++i) {
            b[nb].n = pb->vdims[i].n;
            b[nb].is = b[nb].os = pb->vdims[i].os;
            ++nb;
        }
        if (plan_c2c_1d(c, n0, pb->dims[0].os, pb->dims[0].os, nb, b, FK_BUF_OUT, 0, FK_BUF_OUT,
                        0, sw, sw, 0))
            return -1;
        return 1;
    }
}

/* Cache-blocked r2c / c2r (rank >= 2, batched): 1 planned, 0 not, -1 error */
static int plan_real_blocked(ctx *c, const fk_problem *pb)
{
    fk_plan *P = c->P;
    size_t cs = csz(P->prec), rs = rsz(P->prec);
    long long Nh = 1, H = 1, k;
    int i, v = chunk_vdim(pb), r = pb->rank;
    if (r < 2 || v < 0) return 0;
    for (i = 0; i < r - 1; ++i) Nh *= pb->dims[i].n;
    Nh *= pb->dims[r - 1].n / 2 + 1;
    for (i = 0; i < pb->vrank; ++i) H *= pb->vdims[i].n;
    k = chunk_units(c, pb->vdims[v].n, (double)Nh * cs * (double)(H / pb->vdims[v].n));
    if (k >= pb->vdims[v].n) return 0;
    desc_add(P, "(blocked %lld of %lld)", k, pb->vdims[v].n);
    if (pb->kind == FK_R2C) return plan_vchunks(c, pb, v, k, chunk_r2c, 0, rs, cs) ? -1 : 1;
    return plan_vchunks(c, pb, v, k, chunk_c2r, 0, cs, rs) ? -1 : 1;
}

static int plan_copy(ctx *c, int vrank, const fk_dim *vdims, int op, size_t esz)
{
    fk_pass proto;
    fk_dim b[FK_MAXRANK];
    int nb;
    memcpy(b, vdims, sizeof(fk_dim) * vrank);
    nb = squeeze(vrank, b);
    nb = merge_dims(nb, b);
    sort_by_is(nb, b);
    memset(&proto, 0, sizeof(proto));
    proto.type = FK_PASS_EW;
    proto.e.op = op;
    proto.e.n = 1;
    proto.e.is = proto.e.os = 0;
    proto.ibuf = FK_BUF_IN;
    proto.obuf = FK_BUF_OUT;
    proto.bytes = 2.0 * (double)esz * (double)batch_count(nb, b);
    emit_loops(c, &proto, nb, b, esz, esz);
    return 0;
}

/* every batch stride of the real side (input: 0, output: 1) even, so the
   real lines can be viewed as complex ones */
static int real_pairs(int nb, const fk_dim *b, int out)
{
    int i;
    for (i = 0; i < nb; ++i)
        if ((out ? b[i].os : b[i].is) % 2) return 0;
    return 1;
}

/* Rank >= 2 real transforms whose last dimension has a cross-row kernel
   (FK_IO_R2C_XR / FK_IO_C2R_XR, X = its rows per tile): dimension r-2
   (n0 = X*m) is split as n = n1 + m*n2 (input rows) / k = X*q + p (output
   rows).  The row pass transforms the X rows n1 + m*n2 of a tile and runs
   the first stage of their column DFT,
       y_p[k] = W_n0^(n1*p) sum_n2 X_(n1+m*n2)[k] W_X^(n2*p),
   into row X*n1 + p; what is left of dimension r-2 is an m-point DFT over
   rows X apart, in place: rows X*q + p (q = 0..m-1) -> X*q + p, so its
   column tiles hold X times more lines than an n0-point column pass could
   and read X times longer runs.  This is the four-step split of
   dft/ct.c:34-45 applied across the rank split of rdft/rank-geq2-rdft2.c
   (rows first, then the other dimensions on n/2+1 columns).  c2r mirrors
   it: the m-point inverse over rows X apart first, in place on the input
   (which multi-dimensional c2r may destroy), then the last inverse stage
   across the tile's rows fused into the c2r row pass.  Returns the kernel
   and the extra batch dim (dims 0..r-3 and the howmany dims merged into at
   most one), or NULL when the problem does not fit. */
static const fk_kernel *xr_kernel(ctx *c, const fk_problem *pb, int io, fk_dim *extra, int *nextra)
{
    int r = pb->rank, i, ne = 0;
    long long nl = pb->dims[r > 0 ? r - 1 : 0].n, n0;
    const fk_kernel *k;
    fk_dim e[FK_MAXRANK * 2];
    if (r < 2 || nl % 2 || pb->in == pb->out || c->r16) return 0;
    /* rows per tile: 8 by default, 4 (half the LDS, a 2x longer column
       pass) as the tuner's alternative */
    if (!find_kernel(c->P->prec, nl / 2, FK_ROW, io)) return 0;
    k = pick_variant(c, c->P->prec, nl / 2, FK_ROW, io);
    if (!k) return 0;
    n0 = pb->dims[r - 2].n;
    if (n0 % k->nl || n0 / k->nl < 2) return 0;
    for (i = 0; i < r - 2; ++i) e[ne++] = pb->dims[i];
    for (i = 0; i < pb->vrank; ++i) e[ne++] = pb->vdims[i];
    ne = squeeze(ne, e);
    ne = merge_dims(ne, e);
    if (ne > 1) return 0;
    if (ne) extra[0] = e[0];
    *nextra = ne;
    return k;
}

static int plan_r2c_xr(ctx *c, const fk_problem *pb, const fk_kernel *kx, const fk_dim *extra,
                       int nextra)
{
    fk_plan *P = c->P;
    int prec = P->prec, r = pb->rank, i, nb, nv = 0;
    size_t cs = csz(prec), rs = rsz(prec), before = (size_t)P->npass;
    const fk_dim *d0 = &pb->dims[r - 2], *dl = &pb->dims[r - 1];
    long long nl = dl->n, nh = nl / 2 + 1, L = nl / 2, X = kx->nl, n0 = d0->n, m = n0 / X, tb = 0;
    fk_dim b[3], lead[FK_MAXRANK], vd[FK_MAXRANK * 2];
    const void *wk = line_table(prec, nl, L + 1, c->st);
    const double *t2 = tw2_table(n0, &tb, c->st);
    unsigned fl = 0;
    int unit;
    if (!wk || !t2) {
        c->err = 1;
        set_err("twiddle table allocation failed");
        return -1;
    }
    /* tile = the X rows n1 + m*n2 of one n1: batch dim 0 must be n2 */
    b[0].n = X;
    b[0].is = m * d0->is;
    b[0].os = d0->os;
    b[1].n = m;
    b[1].is = d0->is;
    b[1].os = X * d0->os;
    nb = 2;
    if (nextra) b[nb++] = extra[0];
    unit = dl->is == 1 && ((uintptr_t)pb->in % (2 * rs)) == 0;
    for (i = 0; i < nb; ++i) unit &= (b[i].is % 2) == 0;
    if (unit) fl |= FK_REAL_UNIT;
    emit_stockham(c, kx, dl->is, dl->os, nb, b, c->bin, 0, c->bout, 0, fl, rs, cs);
    for (; before < (size_t)P->npass; ++before) {
        fk_pass *p = P->pass + before;
        p->a.aux = wk;
        p->a.tlo = t2;
        p->a.thi = t2 + 2 * tb;
        p->a.tN = n0;
        p->a.tb = tb;
        p->bytes = (double)batch_count(nb, b) * ((double)nl * rs + (double)nh * cs);
    }
    if (unit) need_align(c, c->bin, (unsigned)(2 * rs));
    desc_add(P, "(r2c-rows+radix-%lld L=%lld)", X, L);
    if (c->err) return -1;
    /* dims 0..r-3 and the m-point rest of dim r-2, in place on the output */
    for (i = 0; i < r - 2; ++i) {
        lead[i].n = pb->dims[i].n;
        lead[i].is = lead[i].os = pb->dims[i].os;
    }
    lead[r - 2].n = m;
    lead[r - 2].is = lead[r - 2].os = X * d0->os;
    vd[nv].n = nh;
    vd[nv].is = vd[nv].os = dl->os;
    ++nv;
    vd[nv].n = X;
    vd[nv].is = vd[nv].os = d0->os;
    ++nv;
    for (i = 0; i < pb->vrank; ++i) {
        vd[nv].n = pb->vdims[i].n;
        vd[nv].is = vd[nv].os = pb->vdims[i].os;
        ++nv;
    }
    return plan_c2c_dims(c, r - 1, lead, nv, vd, c->bout, c->bout, 0);
}

static int plan_c2r_xr(ctx *c, const fk_problem *pb, const fk_kernel *kx, const fk_dim *extra,
                       int nextra)
{
    fk_plan *P = c->P;
    int prec = P->prec, r = pb->rank, i, nb, nv = 0;
    size_t cs = csz(prec), rs = rsz(prec), before;
    const fk_dim *d0 = &pb->dims[r - 2], *dl = &pb->dims[r - 1];
    long long nl = dl->n, nh = nl / 2 + 1, L = nl / 2, X = kx->nl, n0 = d0->n, m = n0 / X, tb = 0;
    fk_dim b[3], lead[FK_MAXRANK], vd[FK_MAXRANK * 2];
    const void *wk = line_table(prec, nl, L + 1, c->st);
    const double *t2 = tw2_table(n0, &tb, c->st);
    unsigned fl = 0;
    int unit;
    if (!wk || !t2) {
        c->err = 1;
        set_err("twiddle table allocation failed");
        return -1;
    }
    /* dims 0..r-3 and the m-point inverse over rows X apart, in place on
       the input (api/plan-many-dft-c2r.c:41-42: it may be destroyed) */
    for (i = 0; i < r - 2; ++i) {
        lead[i].n = pb->dims[i].n;
        lead[i].is = lead[i].os = pb->dims[i].is;
    }
    lead[r - 2].n = m;
    lead[r - 2].is = lead[r - 2].os = X * d0->is;
    vd[nv].n = nh;
    vd[nv].is = vd[nv].os = dl->is;
    ++nv;
    vd[nv].n = X;
    vd[nv].is = vd[nv].os = d0->is;
    ++nv;
    for (i = 0; i < pb->vrank; ++i) {
        vd[nv].n = pb->vdims[i].n;
        vd[nv].is = vd[nv].os = pb->vdims[i].is;
        ++nv;
    }
    if (plan_c2c_dims(c, r - 1, lead, nv, vd, c->bin, c->bin, 1)) return -1;
    /* tile = the X input rows X*n1 + p of one n1 (batch dim 0 = p); line
       n2 of the tile is output row n1 + m*n2 */
    b[0].n = X;
    b[0].is = d0->is;
    b[0].os = m * d0->os;
    b[1].n = m;
    b[1].is = X * d0->is;
    b[1].os = d0->os;
    nb = 2;
    if (nextra) b[nb++] = extra[0];
    unit = dl->os == 1 && ((uintptr_t)pb->out % (2 * rs)) == 0;
    for (i = 0; i < nb; ++i) unit &= (b[i].os % 2) == 0;
    if (unit) fl |= FK_REAL_UNIT;
    before = (size_t)P->npass;
    emit_stockham(c, kx, dl->is, dl->os, nb, b, c->bin, 0, c->bout, 0, fl, cs, rs);
    for (; before < (size_t)P->npass; ++before) {
        fk_pass *p = P->pass + before;
        p->a.aux = wk;
        p->a.tlo = t2;
        p->a.thi = t2 + 2 * tb;
        p->a.tN = n0;
        p->a.tb = tb;
        p->bytes = (double)batch_count(nb, b) * ((double)nl * rs + (double)nh * cs);
    }
    if (unit) need_align(c, c->bout, (unsigned)(2 * rs));
    desc_add(P, "(radix-%lld+c2r-rows L=%lld)", X, L);
    return c->err ? -1 : 0;
}

static int plan_r2c(ctx *c, const fk_problem *pb)
{
    fk_plan *P = c->P;
    int prec = P->prec, r = pb->rank, i;
    size_t cs = csz(prec), rs = rsz(prec);
    long long nl = pb->dims[r - 1].n, nh = nl / 2 + 1, L = nl / 2;
    fk_dim b[FK_MAXRANK * 2], lead[FK_MAXRANK];
    int nb = 0;
    const fk_kernel *k = 0;
    /* batch for the last-dim pass: leading dims + vdims */
    for (i = 0; i < r - 1; ++i) b[nb++] = pb->dims[i];
    for (i = 0; i < pb->vrank; ++i) b[nb++] = pb->vdims[i];
    nb = squeeze(nb, b);
    sort_by_is(nb, b);
    {
        fk_dim xe[1];
        int nxe = 0;
        const fk_kernel *kx = xr_kernel(c, pb, FK_IO_R2C_XR, xe, &nxe);
        if (kx && decide(c, 2, DEC_VARIANT) == 0) return plan_r2c_xr(c, pb, kx, xe, nxe);
    }
    if (nl % 2 == 0) k = pick_variant(c, prec, L, FK_ROW, FK_IO_R2C);
    if (k) {
        size_t before = (size_t)P->npass;
        unsigned fl = 0;
        const void *wk = line_table(prec, nl, L + 1, c->st);
        int unit = pb->dims[r - 1].is == 1 && ((uintptr_t)pb->in % (2 * rs)) == 0;
        for (i = 0; i < nb; ++i) unit &= (b[i].is % 2) == 0;
        if (unit) fl |= FK_REAL_UNIT;
        if (!wk) { c->err = 1; return -1; }
        emit_stockham(c, k, pb->dims[r - 1].is, pb->dims[r - 1].os, nb, b, c->bin, 0,
                      c->bout, 0, fl, rs, cs);
        for (; before < (size_t)P->npass; ++before) {
            fk_pass *p = P->pass + before;
            p->a.aux = wk;
            p->bytes = (double)batch_count(nb, b) * ((double)nl * rs + (double)nh * cs);
        }
        if (unit) need_align(c, c->bin, (unsigned)(2 * rs));
        desc_add(P, "(r2c-fused L=%lld)", L);
    } else if ((k = pick_variant(c, prec, nl, FK_ROW, FK_IO_R2C_FULL)) != 0) {
        /* one pass: full-length DFT of the real line, half the bins stored */
        size_t before = (size_t)P->npass;
        emit_stockham(c, k, pb->dims[r - 1].is, pb->dims[r - 1].os, nb, b, c->bin, 0, c->bout, 0,
                      0, rs, cs);
        for (; before < (size_t)P->npass; ++before)
            P->pass[before].bytes = (double)batch_count(nb, b) * ((double)nl * rs + (double)nh * cs);
        desc_add(P, "(r2c-full n=%lld)", nl);
    } else if (nl % 2 == 0 && nl >= 4 && pb->dims[r - 1].is == 1 &&
               ((uintptr_t)pb->in % (2 * rs)) == 0 && real_pairs(nb, b, 0)) {
        /* even n without a half-length kernel: the line read as L complex
           z_j = x_2j + i x_2j+1 (no copy), the engine's L-point DFT into
           staging, then X[k] = E[k] + W_n^k O[k] for k <= L */
        long long H = batch_count(nb, b), s = L;
        fk_dim cb[FK_MAXRANK * 2], pbat[FK_MAXRANK * 2];
        fk_pass proto;
        const void *wk = line_table(prec, nl, L + 1, c->st);
        if (!wk) { c->err = 1; return -1; }
        need_scratch(c, 6, cs * (size_t)L * (size_t)H);
        for (i = 0; i < nb; ++i) {
            cb[i].n = pbat[i].n = b[i].n;
            cb[i].is = b[i].is / 2;
            cb[i].os = pbat[i].is = s;
            pbat[i].os = b[i].os;
            s *= b[i].n;
        }
        if (plan_c2c_1d(c, L, 1, 1, nb, cb, c->bin, 0, FK_BUF_S0 + 6, 0, 0, 0, 0)) return -1;
        memset(&proto, 0, sizeof(proto));
        proto.type = FK_PASS_EW;
        proto.e.op = FK_EW_R2C_POST;
        proto.e.n = L + 1;
        proto.e.p0 = L;
        proto.e.is = 1;
        proto.e.os = pb->dims[r - 1].os;
        proto.e.aux = wk;
        proto.ibuf = FK_BUF_S0 + 6;
        proto.obuf = c->bout;
        proto.bytes = (double)H * ((double)L + (double)nh) * cs;
        emit_loops(c, &proto, nb, pbat, cs, cs);
        need_align(c, c->bin, (unsigned)(2 * rs));
        desc_add(P, "(r2c-half n=%lld)", nl);
    } else {
        /* widen -> full complex DFT along the last dim -> copy n/2+1 bins */
        long long H = batch_count(nb, b), s;
        fk_dim sb[FK_MAXRANK * 2];
        fk_pass proto;
        need_scratch(c, 6, cs * (size_t)nl * (size_t)H);
        need_scratch(c, 7, cs * (size_t)nl * (size_t)H);
        s = nl;
        for (i = 0; i < nb; ++i) {
            sb[i].n = b[i].n;
            sb[i].is = b[i].is;
            sb[i].os = s;
            s *= b[i].n;
        }
        memset(&proto, 0, sizeof(proto));
        proto.type = FK_PASS_EW;
        proto.e.op = FK_EW_WIDEN;
        proto.e.n = nl;
        proto.e.is = pb->dims[r - 1].is;
        proto.e.os = 1;
        proto.ibuf = c->bin;
        proto.obuf = FK_BUF_S0 + 6;
        proto.bytes = (double)H * nl * (rs + cs);
        emit_loops(c, &proto, nb, sb, rs, cs);
        for (i = 0; i < nb; ++i) sb[i].is = sb[i].os;
        if (plan_c2c_1d(c, nl, 1, 1, nb, sb, FK_BUF_S0 + 6, 0, FK_BUF_S0 + 7, 0, 0, 0, 0))
            return -1;
        for (i = 0; i < nb; ++i) sb[i].os = b[i].os;
        memset(&proto, 0, sizeof(proto));
        proto.type = FK_PASS_EW;
        proto.e.op = FK_EW_COPY;
        proto.e.n = nh;
        proto.e.is = 1;
        proto.e.os = pb->dims[r - 1].os;
        proto.ibuf = FK_BUF_S0 + 7;
        proto.obuf = c->bout;
        proto.bytes = (double)H * nh * 2 * cs;
        emit_loops(c, &proto, nb, sb, cs, cs);
        desc_add(P, "(r2c-generic n=%lld)", nl);
    }
    if (c->err) return -1;
    /* leading dims: c2c forward in place on the output, nh columns */
    if (r > 1) {
        fk_dim vd[FK_MAXRANK * 2];
        int nv = 0;
        for (i = 0; i < r - 1; ++i) {
            lead[i].n = pb->dims[i].n;
            lead[i].is = lead[i].os = pb->dims[i].os;
        }
        vd[nv].n = nh;
        vd[nv].is = vd[nv].os = pb->dims[r - 1].os;
        ++nv;
        for (i = 0; i < pb->vrank; ++i) {
            vd[nv].n = pb->vdims[i].n;
            vd[nv].is = vd[nv].os = pb->vdims[i].os;
            ++nv;
        }
        if (plan_c2c_dims(c, r - 1, lead, nv, vd, c->bout, c->bout, 0)) return -1;
    }
    return 0;
}

static int plan_c2r(ctx *c, const fk_problem *pb)
{
    fk_plan *P = c->P;
    int prec = P->prec, r = pb->rank, i;
    size_t cs = csz(prec), rs = rsz(prec);
    long long nl = pb->dims[r - 1].n, nh = nl / 2 + 1, L = nl / 2;
    fk_dim b[FK_MAXRANK * 2], lead[FK_MAXRANK];
    int nb = 0;
    const fk_kernel *k = 0;
    {
        fk_dim xe[1];
        int nxe = 0;
        const fk_kernel *kx = xr_kernel(c, pb, FK_IO_C2R_XR, xe, &nxe);
        if (kx && decide(c, 2, DEC_VARIANT) == 0) return plan_c2r_xr(c, pb, kx, xe, nxe);
    }
    /* leading dims first: backward c2c in place on the input (destroys it,
       api/plan-many-dft-c2r.c:41-42) */
    if (r > 1) {
        fk_dim vd[FK_MAXRANK * 2];
        int nv = 0;
        for (i = 0; i < r - 1; ++i) {
            lead[i].n = pb->dims[i].n;
            lead[i].is = lead[i].os = pb->dims[i].is;
        }
        vd[nv].n = nh;
        vd[nv].is = vd[nv].os = pb->dims[r - 1].is;
        ++nv;
        for (i = 0; i < pb->vrank; ++i) {
            vd[nv].n = pb->vdims[i].n;
            vd[nv].is = vd[nv].os = pb->vdims[i].is;
            ++nv;
        }
        if (plan_c2c_dims(c, r - 1, lead, nv, vd, c->bin, c->bin, 1)) return -1;
    }
    for (i = 0; i < r - 1; ++i) b[nb++] = pb->dims[i];
    for (i = 0; i < pb->vrank; ++i) b[nb++] = pb->vdims[i];
    nb = squeeze(nb, b);
    sort_by_is(nb, b);
    if (nl % 2 == 0) k = pick_variant(c, prec, L, FK_ROW, FK_IO_C2R);
    if (k) {
        size_t before = (size_t)P->npass;
        unsigned fl = 0;
        const void *wk = line_table(prec, nl, L + 1, c->st);
        int unit = pb->dims[r - 1].os == 1 && ((uintptr_t)pb->out % (2 * rs)) == 0;
        for (i = 0; i < nb; ++i) unit &= (b[i].os % 2) == 0;
        if (unit) fl |= FK_REAL_UNIT;
        if (!wk) { c->err = 1; return -1; }
        emit_stockham(c, k, pb->dims[r - 1].is, pb->dims[r - 1].os, nb, b, c->bin, 0,
                      c->bout, 0, fl, cs, rs);
        for (; before < (size_t)P->npass; ++before) {
            fk_pass *p = P->pass + before;
            p->a.aux = wk;
            p->bytes = (double)batch_count(nb, b) * ((double)nl * rs + (double)nh * cs);
        }
        if (unit) need_align(c, c->bout, (unsigned)(2 * rs));
        desc_add(P, "(c2r-fused L=%lld)", L);
    } else if ((k = pick_variant(c, prec, nl, FK_ROW, FK_IO_C2R_FULL)) != 0) {
        /* one pass: Hermitian expansion on load, real part on store */
        size_t before = (size_t)P->npass;
        emit_stockham(c, k, pb->dims[r - 1].is, pb->dims[r - 1].os, nb, b, c->bin, 0, c->bout, 0,
                      0, cs, rs);
        for (; before < (size_t)P->npass; ++before)
            P->pass[before].bytes = (double)batch_count(nb, b) * ((double)nl * rs + (double)nh * cs);
        desc_add(P, "(c2r-full n=%lld)", nl);
    } else if (nl % 2 == 0 && nl >= 4 && pb->dims[r - 1].os == 1 &&
               ((uintptr_t)pb->out % (2 * rs)) == 0 && real_pairs(nb, b, 1)) {
        /* even n without a half-length kernel: Z[k] = E[k] + i O[k] into
           staging, then the engine's L-point inverse DFT written straight to
           the output viewed as L complex (y_2m = Re z_m, y_2m+1 = Im z_m) */
        long long H = batch_count(nb, b), s = L;
        fk_dim pbat[FK_MAXRANK * 2], cb[FK_MAXRANK * 2];
        fk_pass proto;
        const void *wk = line_table(prec, nl, L + 1, c->st);
        if (!wk) { c->err = 1; return -1; }
        need_scratch(c, 6, cs * (size_t)L * (size_t)H);
        for (i = 0; i < nb; ++i) {
            pbat[i].n = cb[i].n = b[i].n;
            pbat[i].is = b[i].is;
            pbat[i].os = cb[i].is = s;
            cb[i].os = b[i].os / 2;
            s *= b[i].n;
        }
        memset(&proto, 0, sizeof(proto));
        proto.type = FK_PASS_EW;
        proto.e.op = FK_EW_C2R_PRE;
        proto.e.n = L;
        proto.e.p0 = L;
        proto.e.is = pb->dims[r - 1].is;
        proto.e.os = 1;
        proto.e.aux = wk;
        proto.ibuf = c->bin;
        proto.obuf = FK_BUF_S0 + 6;
        proto.bytes = (double)H * ((double)nh + (double)L) * cs;
        emit_loops(c, &proto, nb, pbat, cs, cs);
        if (plan_c2c_1d(c, L, 1, 1, nb, cb, FK_BUF_S0 + 6, 0, c->bout, 0, 1, 1, 0)) return -1;
        need_align(c, c->bout, (unsigned)(2 * rs));
        desc_add(P, "(c2r-half n=%lld)", nl);
    } else {
        /* Hermitian expansion -> full complex inverse -> real part */
        long long H = batch_count(nb, b), s;
        fk_dim sb[FK_MAXRANK * 2];
        fk_pass proto;
        need_scratch(c, 6, cs * (size_t)nl * (size_t)H);
        need_scratch(c, 7, cs * (size_t)nl * (size_t)H);
        s = nl;
        for (i = 0; i < nb; ++i) {
            sb[i].n = b[i].n;
            sb[i].is = b[i].is;
            sb[i].os = s;
            s *= b[i].n;
        }
        memset(&proto, 0, sizeof(proto));
        proto.type = FK_PASS_EW;
        proto.e.op = FK_EW_HERM;
        proto.e.n = nl;
        proto.e.p0 = nl;
        proto.e.is = pb->dims[r - 1].is;
        proto.e.os = 1;
        proto.ibuf = c->bin;
        proto.obuf = FK_BUF_S0 + 6;
        proto.bytes = (double)H * (nh + nl) * cs;
        emit_loops(c, &proto, nb, sb, cs, cs);
        for (i = 0; i < nb; ++i) sb[i].is = sb[i].os;
        if (plan_c2c_1d(c, nl, 1, 1, nb, sb, FK_BUF_S0 + 6, 0, FK_BUF_S0 + 7, 0, 1, 1, 0))
            return -1;
        for (i = 0; i < nb; ++i) sb[i].os = b[i].os;
        memset(&proto, 0, sizeof(proto));
        proto.type = FK_PASS_EW;
        proto.e.op = FK_EW_REAL;
        proto.e.n = nl;
        proto.e.is = 1;
        proto.e.os = pb->dims[r - 1].os;
        proto.ibuf = FK_BUF_S0 + 7;
        proto.obuf = c->bout;
        proto.bytes = (double)H * nl * (cs + rs);
        emit_loops(c, &proto, nb, sb, cs, rs);
        desc_add(P, "(c2r-generic n=%lld)", nl);
    }
    return c->err ? -1 : 0;
}

/* ------------------------------------------------------------------ */
/* real-to-real kinds                                                  */

/* Each dimension of an r2r problem is one pass group: a pre pass from the
   source into a contiguous staging line, an r2c / c2r / c2c of that line,
   and a post pass into the destination with its strides.  What each kind
   computes is doc/reference.texi:2021-2350 (DCT-I..IV, DST-I..IV, DHT) and
   :913-930 (halfcomplex); the staging algorithms are:
     R2HC, DHT       n-point r2c straight from the source, then the
                     halfcomplex / Hartley post pass;
     HC2R            Hermitian bins from halfcomplex, n-point c2r straight
                     into the destination;
     REDFT10/RODFT10 reordering (Makhoul), n-point r2c, post twiddle
                     exp(-i pi k / 2n) (FFTW: reodft/reodft010e-r2hc.c);
     REDFT01/RODFT01 pre twiddle to n/2+1 bins, n-point c2r, reordering;
     REDFT11/RODFT11 even n: n/2-point c2c of (x_2m + i x_{n-1-2m}) with pre
                     and post twiddles; odd n: 2n-point c2c of the
                     twiddled, zero-padded line;
     REDFT00/RODFT00 r2c of the even / odd extension of logical size
                     2(n-1) / 2(n+1) (reodft/redft00e-r2hc-pad.c,
                     reodft/rodft00e-r2hc-pad.c).
   The sine kinds reverse or negate samples around the cosine algorithms. */

static int r2r_trivial(int kind, long long n)
{
    /* rdft/problem.c:129-133: size-1 dims of these kinds are the identity */
    return n == 1 && (kind == FK_K_R2HC || kind == FK_K_HC2R || kind == FK_K_DHT ||
                      kind == FK_K_REDFT01 || kind == FK_K_RODFT01);
}

static const char *r2r_name(int kind)
{
    static const char *nm[] = {"r2hc", "hc2r", "dht", "e00", "e01", "e10", "e11",
                               "o00", "o01", "o10", "o11"};
    return kind >= 0 && kind <= FK_K_RODFT11 ? nm[kind] : "?";
}

static void emit_r2r_ew(ctx *c, int op, long long n, int sine, long long cnt, long long is,
                        long long os, int nb, const fk_dim *b, int ib, int ob, size_t isz,
                        size_t osz, const void *wt, double bytes)
{
    fk_pass proto;
    memset(&proto, 0, sizeof(proto));
    proto.type = FK_PASS_EW;
    proto.e.op = op;
    proto.e.n = cnt;
    proto.e.is = is;
    proto.e.os = os;
    proto.e.p0 = n;
    proto.e.p1 = sine;
    proto.e.aux = wt;
    proto.ibuf = ib;
    proto.obuf = ob;
    proto.bytes = bytes;
    emit_loops(c, &proto, nb, b, isz, osz);
}

/* rank-1 r2c (kind FK_R2C) or c2r over a batch, between buffers ib / ob */
static int plan_real_line(ctx *c, int kind, long long n, long long is, long long os, int nb,
                          const fk_dim *b, int ib, void *iptr, int ob, void *optr)
{
    fk_problem q;
    int sbin = c->bin, sbout = c->bout, rc;
    memset(&q, 0, sizeof(q));
    q.kind = kind;
    q.prec = c->P->prec;
    q.sign = kind == FK_R2C ? -1 : 1;
    q.rank = 1;
    q.dims[0].n = n;
    q.dims[0].is = is;
    q.dims[0].os = os;
    q.vrank = nb;
    memcpy(q.vdims, b, sizeof(fk_dim) * nb);
    q.in = iptr;
    q.out = optr;
    c->bin = ib;
    c->bout = ob;
    rc = kind == FK_R2C ? plan_r2c(c, &q) : plan_c2r(c, &q);
    c->bin = sbin;
    c->bout = sbout;
    return rc;
}

static int plan_r2r_1d(ctx *c, int kind, long long n, long long is, long long os, int nb0,
                       const fk_dim *b0, int ib, void *iptr, int ob, void *optr)
{
    fk_plan *P = c->P;
    int prec = P->prec, nb, i, sine, pre = -1, post = -1, inner, acplx, bcplx;
    size_t cs = csz(prec), rs = rsz(prec), as, bs;
    fk_dim b[FK_MAXRANK * 2], pa[FK_MAXRANK * 2], in_[FK_MAXRANK * 2], pb_[FK_MAXRANK * 2];
    long long H, la = 0, lb = 0, M = n, sa, sb;
    const void *wt = 0;
    const int SA = FK_BUF_S0 + FK_R2R_SLOT, SB = FK_BUF_S0 + FK_R2R_SLOT + 1;
    memcpy(b, b0, sizeof(fk_dim) * nb0);
    nb = squeeze(nb0, b);
    sort_by_is(nb, b);
    H = batch_count(nb, b);
    sine = kind >= FK_K_RODFT00;
    if (n == 1) {
        /* size-1 dims with a constant factor: REDFT10 / RODFT10 / RODFT00
           give 2 x_0, REDFT11 / RODFT11 give sqrt(2) x_0 */
        fk_pass proto;
        memset(&proto, 0, sizeof(proto));
        proto.type = FK_PASS_EW;
        proto.e.op = FK_EW_COPY_R;
        proto.e.n = 1;
        proto.e.is = is;
        proto.e.os = os;
        proto.e.flags = FK_SCALE;
        proto.e.scale = (kind == FK_K_REDFT11 || kind == FK_K_RODFT11) ? sqrt(2.0) : 2.0;
        proto.ibuf = ib;
        proto.obuf = ob;
        proto.bytes = 2.0 * rs * (double)H;
        emit_loops(c, &proto, nb, b, rs, rs);
        return c->err ? -1 : 0;
    }
    {
        /* one pass: the kind's reordering, extension and twiddles fused
           into a real kernel of half length L (R2C for R2HC / DHT /
           REDFT10 / RODFT10 with L = n/2, REDFT00 with L = n-1, RODFT00
           with L = n+1; C2R for HC2R / REDFT01 / RODFT01 with L = n/2) */
        static const int rmode[] = {FK_RM_HC, FK_RM_HC2R, FK_RM_DHT, FK_RM_E00, FK_RM_DCT3,
                                    FK_RM_DCT2, FK_RM_DCT4, FK_RM_O00, FK_RM_DST3, FK_RM_DST2,
                                    FK_RM_DST4};
        const int rm = rmode[kind];
        const int io = (rm == FK_RM_DCT4 || rm == FK_RM_DST4) ? FK_IO_DCT4
                       : (rm == FK_RM_HC2R || rm == FK_RM_DCT3 || rm == FK_RM_DST3) ? FK_IO_C2R_RR
                                                                                  : FK_IO_R2C_RR;
        const long long L = rm == FK_RM_E00 ? n - 1 : rm == FK_RM_O00 ? n + 1 : n / 2;
        const int ok = rm && (rm == FK_RM_E00 || rm == FK_RM_O00 || n % 2 == 0);
        const fk_kernel *k = 0;
        int dC = -1;
        if (ok && (llabs_(is) != 1 || llabs_(os) != 1)) {
            /* strided lines: the column mapping over a unit-stride batch dim */
            long long best = 1LL << 62;
            for (i = 0; i < nb; ++i) {
                long long sc = llabs_(b[i].is) + llabs_(b[i].os);
                if (sc < best && b[i].n >= 2) {
                    best = sc;
                    dC = i;
                }
            }
            if (best > 4) dC = -1;
            if (dC >= 0) k = pick_variant(c, prec, L, FK_COL, io);
            if (k) {
                fk_dim t = b[dC];
                memmove(b + 1, b, sizeof(fk_dim) * dC);
                b[0] = t;
            }
        }
        if (ok && !k) k = pick_variant(c, prec, L, FK_ROW, io);
        if (k) {
            size_t before = (size_t)P->npass;
            unsigned fl = 0;
            int unit = (rm == FK_RM_HC || rm == FK_RM_DHT || rm == FK_RM_HC2R) &&
                       (io == FK_IO_R2C_RR ? is : os) == 1 &&
                       ((uintptr_t)(io == FK_IO_R2C_RR ? iptr : optr) % (2 * rs)) == 0;
            const void *wk = line_table(prec, 2 * L, L + 1, c->st);
            for (i = 0; i < nb; ++i) unit &= ((io == FK_IO_R2C_RR ? b[i].is : b[i].os) % 2) == 0;
            const int twid = rm == FK_RM_DCT2 || rm == FK_RM_DST2 || rm == FK_RM_DCT3 ||
                             rm == FK_RM_DST3 || rm == FK_RM_DCT4 || rm == FK_RM_DST4;
            if (twid) wt = line_table(prec, 8 * n, 2 * n, c->st);
            if (!wk || (twid && !wt)) {
                c->err = 1;
                set_err("twiddle table allocation failed");
                return -1;
            }
            if (unit) fl |= FK_REAL_UNIT;
            if (prec == 1 && k->mode == FK_ROW && (rm == FK_RM_DCT2 || rm == FK_RM_DST2) && is == 1 &&
                ((uintptr_t)iptr % 16) == 0) {
                int quad = 1;
                for (i = 0; i < nb; ++i) quad &= (b[i].is % 4) == 0;
                if (quad) {
                    fl |= FK_REAL_QUAD;
                    need_align(c, ib, 16);
                }
            }
            emit_stockham(c, k, is, os, nb, b, ib, 0, ob, 0, fl, rs, rs);
            for (; before < (size_t)P->npass; ++before) {
                fk_pass *p = P->pass + before;
                p->a.aux = wk;
                p->a.rtw = wt;
                p->a.rmode = (unsigned)rm;
                p->bytes = 2.0 * (double)batch_count(nb, b) * (double)n * rs;
            }
            if (unit) need_align(c, io == FK_IO_R2C_RR ? ib : ob, (unsigned)(2 * rs));
            return c->err ? -1 : 0;
        }
    }
    switch (kind) {
    case FK_K_R2HC:
    case FK_K_DHT:
        inner = FK_R2C; post = FK_EW_HC_POST; sine = kind == FK_K_DHT;
        acplx = 0; bcplx = 1; lb = n / 2 + 1;
        break;
    case FK_K_HC2R:
        inner = FK_C2R; pre = FK_EW_HC_PRE;
        acplx = 1; bcplx = 0; la = n / 2 + 1;
        break;
    case FK_K_REDFT10:
    case FK_K_RODFT10:
        inner = FK_R2C; pre = FK_EW_DCT2_PRE; post = FK_EW_DCT2_POST;
        acplx = 0; bcplx = 1; la = n; lb = n / 2 + 1;
        break;
    case FK_K_REDFT01:
    case FK_K_RODFT01:
        inner = FK_C2R; pre = FK_EW_DCT3_PRE; post = FK_EW_DCT3_POST;
        acplx = 1; bcplx = 0; la = n / 2 + 1; lb = n;
        break;
    case FK_K_REDFT11:
    case FK_K_RODFT11:
        inner = FK_C2C; acplx = bcplx = 1;
        if (n % 2 == 0) {
            pre = FK_EW_DCT4_PRE; post = FK_EW_DCT4_POST; M = la = lb = n / 2;
        } else {
            pre = FK_EW_DCT4O_PRE; post = FK_EW_DCT4O_POST; M = la = lb = 2 * n;
        }
        break;
    case FK_K_REDFT00:
    case FK_K_RODFT00:
        inner = FK_R2C; pre = FK_EW_SYM_PRE; post = FK_EW_SYM_POST;
        M = la = kind == FK_K_REDFT00 ? 2 * (n - 1) : 2 * (n + 1);
        acplx = 0; bcplx = 1; lb = M / 2 + 1;
        break;
    default:
        set_err("invalid r2r kind %d", kind);
        return -1;
    }
    as = acplx ? cs : rs;
    bs = bcplx ? cs : rs;
    if (pre == FK_EW_DCT2_PRE || pre == FK_EW_DCT3_PRE || pre == FK_EW_DCT4_PRE ||
        pre == FK_EW_DCT4O_PRE) {
        wt = line_table(prec, 8 * n, 2 * n, c->st);
        if (!wt) { c->err = 1; set_err("twiddle table allocation failed"); return -1; }
    }
    /* staging layouts: lines contiguous, batch dims in the order of b */
    sa = la;
    sb = lb;
    for (i = 0; i < nb; ++i) {
        pa[i].n = in_[i].n = pb_[i].n = b[i].n;
        pa[i].is = b[i].is;
        pa[i].os = sa;
        in_[i].is = sa;
        in_[i].os = sb;
        pb_[i].is = sb;
        pb_[i].os = b[i].os;
        sa *= b[i].n;
        sb *= b[i].n;
    }
    if (pre >= 0) {
        need_scratch(c, FK_R2R_SLOT, as * (size_t)la * (size_t)H);
        emit_r2r_ew(c, pre, n, sine, la, is, 1, nb, pa, ib, SA, rs, as, wt,
                    (double)H * ((double)n * rs + (double)la * as));
    } else {
        /* the inner transform reads the source directly */
        for (i = 0; i < nb; ++i) in_[i].is = b[i].is;
    }
    if (post >= 0) {
        need_scratch(c, FK_R2R_SLOT + 1, bs * (size_t)lb * (size_t)H);
    } else {
        for (i = 0; i < nb; ++i) in_[i].os = b[i].os;
    }
    if (c->err) return -1;
    if (inner == FK_C2C) {
        if (plan_c2c_1d(c, M, 1, 1, nb, in_, SA, 0, SB, 0, 0, 0, 0)) return -1;
    } else {
        int iib = pre >= 0 ? SA : ib, oob = post >= 0 ? SB : ob;
        if (plan_real_line(c, inner, M, pre >= 0 ? 1 : is, post >= 0 ? 1 : os, nb, in_, iib,
                           pre >= 0 ? 0 : iptr, oob, post >= 0 ? 0 : optr))
            return -1;
    }
    if (post >= 0)
        emit_r2r_ew(c, post, n, sine, n, 1, os, nb, pb_, SB, ob, bs, rs, wt,
                    (double)H * ((double)lb * bs + (double)n * rs));
    return c->err ? -1 : 0;
}

/* Multi-dimensional r2r: the separable product of the 1-D kinds
   (doc/reference.texi:2352-2358), one dimension after another from the
   last; the first group reads the input, the others work in place on the
   output. */
static int plan_r2r(ctx *c, const fk_problem *pb)
{
    fk_plan *P = c->P;
    fk_dim d[FK_MAXRANK];
    int kd[FK_MAXRANK], r = 0, i, s;
    for (i = 0; i < pb->rank; ++i) {
        if (pb->r2r[i] < 0 || pb->r2r[i] > FK_K_RODFT11) {
            set_err("invalid r2r kind %d", pb->r2r[i]);
            return -1;
        }
        if (pb->r2r[i] == FK_K_REDFT00 && pb->dims[i].n == 1) {
            /* doc/reference.texi:2083: not defined for n = 1 */
            set_err("REDFT00 of size 1 is not defined");
            return -1;
        }
    }
    desc_add(P, "(r2r rank=%d", pb->rank);
    for (i = 0; i < pb->rank; ++i) desc_add(P, " %s%lld", r2r_name(pb->r2r[i]), pb->dims[i].n);
    desc_add(P, ")");
    for (i = 0; i < pb->rank; ++i) {
        if (r2r_trivial(pb->r2r[i], pb->dims[i].n)) continue;
        d[r] = pb->dims[i];
        kd[r++] = pb->r2r[i];
    }
    if (r == 0) {
        if (!P->inplace) plan_copy(c, pb->vrank, pb->vdims, FK_EW_COPY_R, rsz(P->prec));
        return c->err ? -1 : 0;
    }
    for (s = 0; s < r; ++s) {
        int dd = r - 1 - s, nb = 0, first = s == 0;
        fk_dim b[FK_MAXRANK * 2];
        for (i = 0; i < r; ++i) {
            if (i == dd) continue;
            b[nb] = d[i];
            if (!first) b[nb].is = b[nb].os;
            ++nb;
        }
        for (i = 0; i < pb->vrank; ++i) {
            b[nb] = pb->vdims[i];
            if (!first) b[nb].is = b[nb].os;
            ++nb;
        }
        if (plan_r2r_1d(c, kd[dd], d[dd].n, first ? d[dd].is : d[dd].os, d[dd].os, nb, b,
                        first ? FK_BUF_IN : FK_BUF_OUT, first ? pb->in : pb->out, FK_BUF_OUT,
                        pb->out))
            return -1;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* validation and normalisation                                        */

static int dims_kosher(const fk_problem *pb)
{
    int i;
    if (pb->rank < 0 || pb->rank > FK_MAXRANK || pb->vrank < 0 || pb->vrank > FK_MAXRANK)
        return 0;
    for (i = 0; i < pb->rank; ++i)
        if (pb->dims[i].n <= 0) return 0;
    for (i = 0; i < pb->vrank; ++i)
        if (pb->vdims[i].n < 0) return 0;
    return 1;
}

static int inplace_ok(const fk_problem *pb)
{
    /* FFTW: in-place needs identical input and output locations
       (dft/problem.c:95-99, kernel/tensor7.c:201-215) */
    int i;
    for (i = 0; i < pb->rank; ++i)
        if (pb->dims[i].n > 1 && pb->dims[i].is != pb->dims[i].os) return 0;
    for (i = 0; i < pb->vrank; ++i)
        if (pb->vdims[i].n > 1 && pb->vdims[i].is != pb->vdims[i].os) return 0;
    return 1;
}

/* Contiguous row-major layout of a complex tensor: dims (n of the complex
   side), then vdims outermost; writes strides into sd / sv, returns the
   element count. */
static long long contig_layout(int rank, const fk_dim *dims, int vrank, const fk_dim *vdims,
                               long long nlast, fk_dim *sd, fk_dim *sv)
{
    long long s = 1;
    int i;
    for (i = rank - 1; i >= 0; --i) {
        sd[i].n = i == rank - 1 ? nlast : dims[i].n;
        sd[i].is = sd[i].os = s;
        s *= sd[i].n;
    }
    for (i = vrank - 1; i >= 0; --i) {
        sv[i].n = vdims[i].n;
        sv[i].is = sv[i].os = s;
        s *= sv[i].n;
    }
    return s;
}

/* Elementwise pass between the split arrays of the user (strides in reals,
   from dims / vdims) and the interleaved staging buffer (layout sd / sv):
   gather = split -> staging, scatter = staging -> split. */
static void emit_split_ew(ctx *c, int gather, int rank, const fk_dim *dims, int vrank,
                          const fk_dim *vdims, const fk_dim *sd, const fk_dim *sv)
{
    fk_pass proto;
    fk_dim b[FK_MAXRANK * 2];
    int nb = 0, i, prec = c->P->prec;
    size_t cs = csz(prec), rs = rsz(prec);
    long long cnt = 1;
    memset(&proto, 0, sizeof(proto));
    proto.type = FK_PASS_EW;
    proto.e.op = gather ? FK_EW_SPLIT_GATHER : FK_EW_SPLIT_SCATTER;
    if (rank > 0) {
        proto.e.n = sd[rank - 1].n;
        proto.e.is = gather ? dims[rank - 1].is : 1;
        proto.e.os = gather ? 1 : dims[rank - 1].os;
    } else {
        proto.e.n = 1;
    }
    for (i = 0; i < rank - 1; ++i, ++nb) {
        b[nb].n = sd[i].n;
        b[nb].is = gather ? dims[i].is : sd[i].is;
        b[nb].os = gather ? sd[i].os : dims[i].os;
    }
    for (i = 0; i < vrank; ++i, ++nb) {
        b[nb].n = sv[i].n;
        b[nb].is = gather ? vdims[i].is : sv[i].is;
        b[nb].os = gather ? sv[i].os : vdims[i].os;
    }
    for (i = 0; i < nb; ++i) cnt *= b[i].n;
    cnt *= proto.e.n;
    if (gather) {
        proto.ibuf = FK_BUF_IN;
        proto.i2buf = FK_BUF_IN2;
        proto.obuf = FK_BUF_S0 + FK_SPLIT_SLOT;
    } else {
        proto.ibuf = FK_BUF_S0 + FK_SPLIT_SLOT;
        proto.obuf = FK_BUF_OUT;
        proto.o2buf = FK_BUF_OUT2;
    }
    proto.bytes = (double)cnt * (double)(cs + 2 * rs);
    emit_loops(c, &proto, nb, b, gather ? rs : cs, gather ? cs : rs);
}

/* Split-complex problems with separate real and imaginary arrays
   (api/plan-guru-split-dft.c, api/plan-guru-split-dft-r2c.c,
   api/plan-guru-split-dft-c2r.c): the complex side is staged through an
   interleaved buffer, so every split plan is gather -> interleaved plan ->
   scatter (c2c), interleaved r2c -> scatter, or gather -> c2r.  The
   transform is the forward one on (ri, ii); a backward transform is the
   caller swapping the real and imaginary pointers, as in FFTW. */
static int plan_split(ctx *c, const fk_problem *pb)
{
    fk_plan *P = c->P;
    int r = pb->rank, i, prec = P->prec;
    fk_dim sd[FK_MAXRANK], sv[FK_MAXRANK];
    long long nlast = r > 0 ? pb->dims[r - 1].n : 1, cnt;
    const int S = FK_BUF_S0 + FK_SPLIT_SLOT;
    if (pb->kind != FK_C2C && r == 0) {
        set_err("rank-0 split real transforms are not supported");
        return -1;
    }
    if (pb->kind != FK_C2C) nlast = nlast / 2 + 1;
    cnt = contig_layout(r, pb->dims, pb->vrank, pb->vdims, nlast, sd, sv);
    need_scratch(c, FK_SPLIT_SLOT, csz(prec) * (size_t)cnt);
    if (pb->kind == FK_C2C) {
        desc_add(P, "(dft-split rank=%d", r);
        for (i = 0; i < r; ++i) desc_add(P, " %lld", pb->dims[i].n);
        desc_add(P, ")");
        emit_split_ew(c, 1, r, pb->dims, pb->vrank, pb->vdims, sd, sv);
        if (r > 0 && plan_c2c_dims(c, r, sd, pb->vrank, sv, S, S, 0)) return -1;
        emit_split_ew(c, 0, r, pb->dims, pb->vrank, pb->vdims, sd, sv);
    } else if (pb->kind == FK_R2C) {
        fk_problem q = *pb;
        for (i = 0; i < r; ++i) q.dims[i].os = sd[i].os;
        for (i = 0; i < pb->vrank; ++i) q.vdims[i].os = sv[i].os;
        q.split = FK_SPLIT_NONE;
        c->bout = S;
        desc_add(P, "(rdft2-split r2c)");
        if (plan_r2c(c, &q)) return -1;
        c->bout = FK_BUF_OUT;
        emit_split_ew(c, 0, r, pb->dims, pb->vrank, pb->vdims, sd, sv);
    } else {
        fk_problem q = *pb;
        /* the staged copy is the c2r input, so the caller's arrays are
           preserved */
        for (i = 0; i < r; ++i) q.dims[i].is = sd[i].is;
        for (i = 0; i < pb->vrank; ++i) q.vdims[i].is = sv[i].is;
        q.split = FK_SPLIT_NONE;
        emit_split_ew(c, 1, r, pb->dims, pb->vrank, pb->vdims, sd, sv);
        c->bin = S;
        desc_add(P, "(rdft2-split c2r)");
        if (plan_c2r(c, &q)) return -1;
        c->bin = FK_BUF_IN;
    }
    return c->err ? -1 : 0;
}

static fk_plan *build_plan(const fk_problem *pb0, hipStream_t st, const int *choice, int nchoice,
                           ctx *cout)
{
    fk_problem pb = *pb0;
    fk_plan *P;
    ctx c;
    int i, ok = 0, rb;
    if (fk_init()) return 0;
    if (!dims_kosher(&pb)) {
        set_err("invalid dimensions");
        return 0;
    }
    P = (fk_plan *)calloc(1, sizeof(fk_plan));
    P->kind = pb.kind;
    P->prec = pb.prec;
    P->sign = pb.sign;
    P->flags = pb.flags;
    P->rank = pb.rank;
    P->vrank = pb.vrank;
    memcpy(P->dims, pb.dims, sizeof(fk_dim) * pb.rank);
    memcpy(P->vdims, pb.vdims, sizeof(fk_dim) * pb.vrank);
    P->in = pb.in;
    P->out = pb.out;
    P->refcount = 1;
    P->stream = g_stream;
    P->align_need = (unsigned)(pb.kind == FK_C2C ? csz(pb.prec) : rsz(pb.prec));
    memset(&c, 0, sizeof(c));
    c.P = P;
    c.st = st;
    c.err = 0;
    c.choice = choice;
    c.nchoice = nchoice;
    c.r16 = (pb.flags & FK_MFMA) && pb.prec == 1 && pb.kind == FK_C2C;
    c.bin = FK_BUF_IN;
    c.bout = FK_BUF_OUT;
    P->split = pb.split;
    P->in2 = pb.in2;
    P->out2 = pb.out2;
    if (pb.split == FK_SPLIT_GENERAL) P->align_need = (unsigned)rsz(pb.prec);
    for (i = 0; i < pb.vrank; ++i)
        if (pb.vdims[i].n == 0) goto done_ok; /* howmany = 0: nothing to do */
    if (pb.split == FK_SPLIT_GENERAL) {
        if (plan_split(&c, &pb)) goto fail;
    } else if (pb.kind == FK_C2C) {
        int sw = pb.sign != -1;
        P->inplace = pb.in == pb.out;
        if (P->inplace && !inplace_ok(&pb)) goto fail;
        if (((uintptr_t)pb.in | (uintptr_t)pb.out) % csz(pb.prec)) {
            set_err("arrays must be aligned to the complex element size");
            goto fail;
        }
        if (pb.rank == 0) {
            if (!P->inplace) plan_copy(&c, pb.vrank, pb.vdims, FK_EW_COPY, csz(pb.prec));
            desc_add(P, "(dft-rank0-copy)");
        } else {
            desc_add(P, "(dft-%s rank=%d", sw ? "bwd" : "fwd", pb.rank);
            for (i = 0; i < pb.rank; ++i) desc_add(P, " %lld", pb.dims[i].n);
            desc_add(P, ")");
            rb = plan_c2c_blocked(&c, &pb, sw);
            if (rb < 0) goto fail;
            if (!rb && plan_c2c_dims(&c, pb.rank, pb.dims, pb.vrank, pb.vdims, FK_BUF_IN,
                                     FK_BUF_OUT, sw))
                goto fail;
        }
    } else if (pb.kind == FK_R2C) {
        P->inplace = pb.in == pb.out;
        if (pb.rank == 0) {
            /* rank-0 rdft2: x -> (x, 0) */
            fk_pass proto;
            fk_dim b[FK_MAXRANK];
            int nb;
            memcpy(b, pb.vdims, sizeof(fk_dim) * pb.vrank);
            nb = squeeze(pb.vrank, b);
            memset(&proto, 0, sizeof(proto));
            proto.type = FK_PASS_EW;
            proto.e.op = FK_EW_WIDEN;
            proto.e.n = 1;
            proto.ibuf = FK_BUF_IN;
            proto.obuf = FK_BUF_OUT;
            emit_loops(&c, &proto, nb, b, rsz(pb.prec), csz(pb.prec));
        } else if ((rb = plan_real_blocked(&c, &pb)) < 0 || (!rb && plan_r2c(&c, &pb))) {
            goto fail;
        }
    } else if (pb.kind == FK_R2R) {
        P->inplace = pb.in == pb.out;
        memcpy(P->r2r, pb.r2r, sizeof(int) * pb.rank);
        if (P->inplace && !inplace_ok(&pb)) goto fail;
        if (((uintptr_t)pb.in | (uintptr_t)pb.out) % rsz(pb.prec)) {
            set_err("arrays must be aligned to the real element size");
            goto fail;
        }
        if (plan_r2r(&c, &pb)) goto fail;
    } else {
        P->inplace = pb.in == pb.out;
        if (pb.rank > 1 && (pb.flags & FK_PRESERVE_INPUT)) {
            /* no input-preserving multi-D c2r, as in FFTW */
            set_err("multi-dimensional c2r cannot preserve its input");
            goto fail;
        }
        if (pb.rank == 0) {
            fk_pass proto;
            fk_dim b[FK_MAXRANK];
            int nb;
            memcpy(b, pb.vdims, sizeof(fk_dim) * pb.vrank);
            nb = squeeze(pb.vrank, b);
            memset(&proto, 0, sizeof(proto));
            proto.type = FK_PASS_EW;
            proto.e.op = FK_EW_REAL;
            proto.e.n = 1;
            proto.ibuf = FK_BUF_IN;
            proto.obuf = FK_BUF_OUT;
            emit_loops(&c, &proto, nb, b, csz(pb.prec), rsz(pb.prec));
        } else if ((rb = plan_real_blocked(&c, &pb)) < 0 || (!rb && plan_c2r(&c, &pb))) {
            goto fail;
        }
    }
    if (c.err) goto fail;
done_ok:
    for (i = 0; i < P->nscratch; ++i) {
        if (P->scratch_bytes[i]) {
            P->scratch[i] = fk_malloc(P->scratch_bytes[i]);
            if (!P->scratch[i]) {
                set_err("scratch allocation of %zu bytes failed", P->scratch_bytes[i]);
                goto fail;
            }
        }
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
        set_err("table generation failed");
        goto fail;
    }
    P->nchoice = c.ci < FK_MAX_DECIDE ? c.ci : FK_MAX_DECIDE;
    if (c.r16) fk_plan_use_mfma(P);
    fk_count_ops(P);
    if (cout) *cout = c;
    ok = 1;
fail:
    if (!ok) {
        fk_plan_destroy(P);
        return 0;
    }
    return P;
}

/* canonical problem key for wisdom (FFTW hashes the problem with md5,
   kernel/planner.c:170-177) */
static unsigned long long problem_hash(const fk_problem *pb)
{
    long long v[8 + 7 * FK_MAXRANK];
    int m = 0, i;
    v[m++] = pb->kind;
    v[m++] = pb->prec;
    v[m++] = pb->kind == FK_C2C ? pb->sign : 0;
    v[m++] = (pb->flags & (FK_PRESERVE_INPUT | FK_DESTROY_INPUT | FK_MFMA | FK_BLOCKED)) |
             ((long long)pb->split << 40);
    v[m++] = pb->in == pb->out;
    v[m++] = (long long)((((uintptr_t)pb->in) | ((uintptr_t)pb->out)) % 16);
    v[m++] = pb->rank;
    v[m++] = pb->vrank;
    for (i = 0; i < pb->rank; ++i) {
        v[m++] = pb->dims[i].n;
        v[m++] = pb->dims[i].is;
        v[m++] = pb->dims[i].os;
        if (pb->kind == FK_R2R) v[m++] = pb->r2r[i];
    }
    for (i = 0; i < pb->vrank; ++i) {
        v[m++] = pb->vdims[i].n;
        v[m++] = pb->vdims[i].is;
        v[m++] = pb->vdims[i].os;
    }
    return fk_fnv1a(v, sizeof(long long) * m, 0);
}

static int patience(unsigned flags)
{
    if (flags & FK_ESTIMATE) return 0;
    if (flags & FK_EXHAUSTIVE) return 3;
    if (flags & FK_PATIENT) return 2;
    return 1; /* FFTW_MEASURE is 0: the default */
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Candidate timing, the method of kernel/timer.c:142-181: `iter`
   back-to-back executes between two device events, the minimum over
   TUNE_REPEAT such runs, with `iter` doubled from 1 until that minimum is
   at least TUNE_MIN_MS of device time; one candidate stops repeating after
   TUNE_LIMIT_S of wall time (FFTW_TIME_LIMIT).  The executes run on the
   planning arrays (FFTW_MEASURE may overwrite them,
   doc/reference.texi:448-450).  Returns ms per execute. */
#define TUNE_REPEAT 8
#define TUNE_MIN_MS 1.0
#define TUNE_LIMIT_S 2.0
static double fk_plan_time_run(fk_plan *p, int iter);
/* The memory-side cache (256 MB on MI355X) would serve a plan whose arrays
   fit in it from the cache on every repeat of the timing loop, which is not
   how the passes meet data in a stream of transforms.  Such a plan is timed
   instead on a rotation of device copies of its arrays (at least twice the
   cache in all), so FFTW_MEASURE ranks candidates on HBM-resident data. */
#define TUNE_CACHE_BYTES ((size_t)256 << 20)
#define TUNE_ROT_MAX 16
typedef struct {
    int n;
    void *buf[2 * TUNE_ROT_MAX];   /* copy i: buf[2i] input, buf[2i + 1] output */
    void *own[2 * TUNE_ROT_MAX];   /* what tune_rot_begin allocated */
} tune_rot;

/* bytes spanned by an array of element size esz addressed by the
   (n, stride) pairs (stride of pick 0: input, 1: output); 0 when a stride
   is negative */
static size_t array_extent(const fk_dim *d, int nd, const fk_dim *v, int nv, int pick, size_t esz)
{
    long long span = 0;
    int i;
    for (i = 0; i < nd + nv; ++i) {
        const fk_dim *x = i < nd ? d + i : v + (i - nd);
        long long st = pick ? x->os : x->is;
        if (st < 0) return 0;
        if (x->n > 1) span += (x->n - 1) * st;
    }
    return (size_t)(span + 1) * esz;
}

static void tune_rot_begin(const fk_problem *pb, tune_rot *r)
{
    size_t ie, oe, isz, osz, foot;
    int i, n;
    memset(r, 0, sizeof(*r));
    r->n = 1;
    r->buf[0] = pb->in;
    r->buf[1] = pb->out;
    if (pb->split != FK_SPLIT_NONE) return;
    isz = pb->kind == FK_R2C || pb->kind == FK_R2R ? rsz(pb->prec) : csz(pb->prec);
    osz = pb->kind == FK_C2R || pb->kind == FK_R2R ? rsz(pb->prec) : csz(pb->prec);
    ie = array_extent(pb->dims, pb->rank, pb->vdims, pb->vrank, 0, isz);
    oe = array_extent(pb->dims, pb->rank, pb->vdims, pb->vrank, 1, osz);
    if (!ie || !oe) return;
    foot = pb->in == pb->out ? (ie > oe ? ie : oe) : ie + oe;
    /* beyond the cache already, or too small for TUNE_ROT_MAX copies to
       leave it: timed on the arrays themselves */
    if (foot >= TUNE_CACHE_BYTES || foot * TUNE_ROT_MAX < 2 * TUNE_CACHE_BYTES) return;
    n = (int)((2 * TUNE_CACHE_BYTES + foot - 1) / foot);
    for (i = 1; i < n; ++i) {
        void *a = 0, *b = 0;
        if (pb->in == pb->out) {
            a = fk_malloc(foot);
            b = a;
        } else {
            a = fk_malloc(ie);
            b = a ? fk_malloc(oe) : 0;
        }
        if (!a || !b) {
            fk_free(a);
            break;
        }
        /* defined contents, so every copy times the same arithmetic */
        (void)hipMemset(a, 0, pb->in == pb->out ? foot : ie);
        if (b != a) (void)hipMemset(b, 0, oe);
        r->own[2 * i] = a;
        r->own[2 * i + 1] = b != a ? b : 0;
        r->buf[2 * i] = a;
        r->buf[2 * i + 1] = b;
        r->n = i + 1;
    }
    (void)hipGetLastError();
    (void)hipDeviceSynchronize();
}

static void tune_rot_end(tune_rot *r)
{
    int i;
    for (i = 0; i < 2 * TUNE_ROT_MAX; ++i) fk_free(r->own[i]);
    memset(r, 0, sizeof(*r));
}

static int cmp_dbl(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : x > y;
}

/* *spread (if given) = (median - minimum) / minimum of the last round's
   repeats: the run-to-run noise of this candidate */
static double time_plan(fk_plan *P, double *spread)
{
    double t0 = now_s(), tmin = 1e30, ts[TUNE_REPEAT];
    int iter, r, nr = 0;
    if (spread) *spread = 0;
    if (fk_plan_time_run(P, 1) < 0) return 1e30; /* warm-up */
    for (iter = 1; iter <= (1 << 16); iter *= 2) {
        tmin = 1e30;
        for (r = nr = 0; r < TUNE_REPEAT; ++r) {
            double t = fk_plan_time_run(P, iter);
            if (t < 0) return 1e30;
            ts[nr++] = t;
            if (t < tmin) tmin = t;
            if (now_s() - t0 > TUNE_LIMIT_S) break;
        }
        if (tmin >= TUNE_MIN_MS || now_s() - t0 > TUNE_LIMIT_S) break;
    }
    if (spread && nr > 1) {
        qsort(ts, (size_t)nr, sizeof(double), cmp_dbl);
        *spread = (ts[nr / 2] - tmin) / tmin;
    }
    return tmin / iter;
}

/* A candidate replaces the current plan only when it is faster by more
   than the larger of 3% and twice the two candidates' combined noise, and a
   second timing of both confirms it: decisions inside the noise keep the
   heuristic choice, so repeated FFTW_MEASURE runs reach the same plan. */
static int clearly_faster(fk_plan *P, double *best, double sp_best, fk_plan *Q, double t,
                          double sp_t)
{
    double thr = 2.0 * (sp_best + sp_t), t2, b2;
    if (thr < 0.03) thr = 0.03;
    if (!(t < *best * (1.0 - thr))) return 0;
    b2 = time_plan(P, 0);
    t2 = time_plan(Q, 0);
    if (!(t2 < b2 * (1.0 - thr))) return 0;
    *best = t < t2 ? t : t2;
    return 1;
}

/*
 * Planner: wisdom lookup, else heuristic plan (ESTIMATE), else greedy
 * search over the planner's decisions timed on the device (MEASURE: kernel
 * geometries and thread mappings; PATIENT/EXHAUSTIVE: also the two-pass
 * splits), the role of kernel/planner.c:518-615 and kernel/timer.c:142-181.
 * The winning decisions become wisdom.
 */
static fk_plan *plan_internal(const fk_problem *pb, hipStream_t st)
{
    int level = patience(pb->flags), wlevel = 0, wn = 0, i, a;
    int wc[FK_MAX_DECIDE], cur[FK_MAX_DECIDE], ncur;
    unsigned long long h = problem_hash(pb);
    fk_plan *P, *Q;
    ctx cb, cq;
    tune_rot rot;
    double best, t0, sp_best = 0, sp_t = 0;
    if (fk_wisdom_lookup(h, &wlevel, wc, &wn) &&
        (wlevel >= level || (pb->flags & FK_WISDOM_ONLY))) {
        P = build_plan(pb, st, wc, wn, 0);
        if (P) P->tuned = 1;
        return P;
    }
    if (pb->flags & FK_WISDOM_ONLY) {
        set_err("FFTW_WISDOM_ONLY: no wisdom for this problem");
        return 0;
    }
    P = build_plan(pb, st, 0, 0, &cb);
    if (!P || level == 0 || !P->in || !P->out || P->npass == 0) return P;
    t0 = now_s();
    tune_rot_begin(pb, &rot);
    P->rot = rot.buf;
    P->nrot = rot.n;
    best = time_plan(P, &sp_best);
    ncur = P->nchoice;
    memcpy(cur, P->choice, sizeof(int) * ncur);
    for (i = 0; i < ncur && i < FK_MAX_DECIDE; ++i) {
        if (cb.dkind[i] > level) continue;
        for (a = 0; a < cb.nalt[i]; ++a) {
            int trial[FK_MAX_DECIDE];
            double t;
            if (a == cur[i]) continue;
            if (g_timelimit >= 0 && now_s() - t0 > g_timelimit) goto done;
            memcpy(trial, cur, sizeof(int) * ncur);
            trial[i] = a;
            Q = build_plan(pb, st, trial, ncur, &cq);
            if (!Q) continue;
            Q->rot = rot.buf;
            Q->nrot = rot.n;
            t = time_plan(Q, &sp_t);
            if (clearly_faster(P, &best, sp_best, Q, t, sp_t)) {
                fk_plan_destroy(P);
                P = Q;
                sp_best = sp_t;
                ncur = Q->nchoice;
                memcpy(cur, Q->choice, sizeof(int) * ncur);
                cb = cq;
            } else {
                fk_plan_destroy(Q);
            }
        }
    }
done:
    P->rot = 0;
    P->nrot = 0;
    tune_rot_end(&rot);
    P->pcost = best * 1e-3;
    P->tuned = 2;
    fk_wisdom_store(h, level, cur, ncur);
    return P;
}

fk_plan *fk_plan_create(const fk_problem *pb)
{
    fk_plan *p;
    pthread_mutex_lock(&g_lock);
    p = plan_internal(pb, g_stream);
    pthread_mutex_unlock(&g_lock);
    return p;
}

void fk_plan_destroy(fk_plan *p)
{
    int i;
    if (!p) return;
    for (i = 0; i < FK_MAX_SCRATCH; ++i) fk_free(p->scratch[i]);
    free(p->pass);
    free(p);
}

/* A plan's own scratch is reused by every execute, and in async mode an
   execute returns once its passes are queued.  Work queued on the old
   stream may still be using the scratch, so the new stream first waits for
   everything queued on the old one: executes of the plan stay ordered
   across the switch. */
void fk_plan_set_stream(fk_plan *p, hipStream_t s)
{
    hipEvent_t ev;
    if (!p || p->stream == s) return;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess) {
        if (hipEventRecord(ev, p->stream) != hipSuccess ||
            hipStreamWaitEvent(s, ev, 0) != hipSuccess) {
            /* cannot order the streams: drain the old one instead */
            (void)hipGetLastError();
            (void)hipStreamSynchronize(p->stream);
        }
        (void)hipEventDestroy(ev);
    } else {
        (void)hipGetLastError();
        (void)hipStreamSynchronize(p->stream);
    }
    p->stream = s;
}

/* ------------------------------------------------------------------ */
/* execution                                                           */

static int launch_pass(const fk_pass *ps, void *const *buf, hipStream_t st, int prec)
{
    if (ps->type == FK_PASS_STOCKHAM) {
        fk_args a = ps->a;
        void *args[1];
        a.in = (const char *)buf[ps->ibuf] + ps->ioff;
        a.out = (char *)buf[ps->obuf] + ps->ooff;
        args[0] = &a;
        return (int)hipLaunchKernel(ps->k->fn, (dim3){ps->grid, 1, 1}, (dim3){ps->block, 1, 1},
                                    args, ps->lds, st);
    } else {
        fk_ew_args e = ps->e;
        e.in = (const char *)buf[ps->ibuf] + ps->ioff;
        e.out = (char *)buf[ps->obuf] + ps->ooff;
        if (e.op == FK_EW_SPLIT_GATHER) e.in2 = (const char *)buf[ps->i2buf] + ps->i2off;
        if (e.op == FK_EW_SPLIT_SCATTER) e.out2 = (char *)buf[ps->o2buf] + ps->o2off;
        return fkm_ew(prec, &e, st);
    }
}

int fk_execute(fk_plan *p, void *in, void *out) { return fk_execute4(p, in, 0, out, 0); }

/* The passes of fk_execute4 with every scratch slot carved (256-byte
   aligned) out of the static device scratch, whose address was taken once
   per device at set-up.  No allocation on any path: a plan whose slots do
   not fit is rejected, and split-array plans (two more arrays) are not
   served.  Synchronizes unless async mode is on. */
int fk_execute_noalloc(fk_plan *p, void *in, void *out)
{
    void *buf[FK_NBUF];
    void *base = 0;
    size_t cap = 0, off = 0;
    int i, rc;
    if (!p) {
        set_err("execute_noalloc: NULL plan");
        return (int)hipErrorInvalidValue;
    }
    if (p->split == FK_SPLIT_GENERAL) {
        set_err("execute_noalloc: split-array plans are not supported");
        return (int)hipErrorInvalidValue;
    }
    if (((uintptr_t)in | (uintptr_t)out) % p->align_need ||
        (p->align_in && ((uintptr_t)in % p->align_in)) ||
        (p->align_out && ((uintptr_t)out % p->align_out))) {
        set_err("execute_noalloc: arrays not aligned as at planning (%u / %u / %u bytes)",
                p->align_need, p->align_in, p->align_out);
        return (int)hipErrorInvalidValue;
    }
    for (i = 0; i < FK_MAX_SCRATCH; ++i) off += (p->scratch_bytes[i] + 255) & ~(size_t)255;
    if (off > FK_STATIC_SCRATCH_BYTES) {
        set_err("execute_noalloc: plan needs %zu bytes of scratch, the static scratch holds %zu",
                off, (size_t)FK_STATIC_SCRATCH_BYTES);
        return (int)hipErrorInvalidValue;
    }
    rc = fk_static_scratch(&base, &cap);
    if (rc || !base) {
        set_err("execute_noalloc: static scratch unavailable (%d)", rc);
        return rc ? rc : (int)hipErrorInvalidSymbol;
    }
    buf[FK_BUF_IN] = in;
    buf[FK_BUF_OUT] = out;
    buf[FK_BUF_IN2] = 0;
    buf[FK_BUF_OUT2] = 0;
    off = 0;
    for (i = 0; i < FK_MAX_SCRATCH; ++i) {
        buf[FK_BUF_S0 + i] = p->scratch_bytes[i] ? (char *)base + off : 0;
        off += (p->scratch_bytes[i] + 255) & ~(size_t)255;
    }
    for (i = 0; i < p->npass && !rc; ++i) rc = launch_pass(p->pass + i, buf, p->stream, p->prec);
    if (!rc && !g_async) rc = (int)hipStreamSynchronize(p->stream);
    if (rc) set_err("execute_noalloc failed: %s", hipGetErrorString((hipError_t)rc));
    return rc;
}

int fk_execute4(fk_plan *p, void *in, void *in2, void *out, void *out2)
{
    void *buf[FK_NBUF];
    void *tmp[FK_MAX_SCRATCH];
    int i, rc = 0, own = 0;
    hipStream_t st = p->stream;
    if (((uintptr_t)in | (uintptr_t)out | (uintptr_t)in2 | (uintptr_t)out2) % p->align_need ||
        (p->align_in && ((uintptr_t)in % p->align_in)) ||
        (p->align_out && ((uintptr_t)out % p->align_out))) {
        set_err("execute: arrays not aligned as at planning (%u / %u / %u bytes)", p->align_need,
                p->align_in, p->align_out);
        return (int)hipErrorInvalidValue;
    }
    if (p->split == FK_SPLIT_GENERAL && p->npass &&
        ((p->kind != FK_R2C && !in2) || (p->kind != FK_C2R && !out2))) {
        set_err("execute: split-array plan needs both real and imaginary arrays");
        return (int)hipErrorInvalidValue;
    }
    buf[FK_BUF_IN] = in;
    buf[FK_BUF_OUT] = out;
    buf[FK_BUF_IN2] = in2;
    buf[FK_BUF_OUT2] = out2;
    /* concurrent executes of one plan each get their own scratch */
    if (__sync_bool_compare_and_swap(&p->scratch_busy, 0, 1)) {
        own = 1;
        for (i = 0; i < FK_MAX_SCRATCH; ++i) buf[FK_BUF_S0 + i] = p->scratch[i];
    } else {
        for (i = 0; i < FK_MAX_SCRATCH; ++i) {
            tmp[i] = 0;
            if (p->scratch_bytes[i] &&
                hipMallocAsync(&tmp[i], p->scratch_bytes[i], st) != hipSuccess)
                rc = (int)hipErrorOutOfMemory;
            buf[FK_BUF_S0 + i] = tmp[i];
        }
    }
    for (i = 0; i < p->npass && !rc; ++i) rc = launch_pass(p->pass + i, buf, st, p->prec);
    if (!own) {
        for (i = 0; i < FK_MAX_SCRATCH; ++i)
            if (tmp[i]) (void)hipFreeAsync(tmp[i], st);
    }
    if (!rc && !g_async) rc = (int)hipStreamSynchronize(st);
    if (own) __sync_lock_release(&p->scratch_busy);
    if (rc) set_err("execute failed: %s", hipGetErrorString((hipError_t)rc));
    return rc;
}

/* Switch every FP32 pass of length 16, 256 or 4096 with all-16 radices to
   the split-precision matrix-core kernel (fk_mfma.hip).  Returns the
   number of passes switched. */
int fkm_mfma_kernel(int L, fk_kernel *k);

static fk_kernel g_mk[3];
static pthread_once_t g_mk_once = PTHREAD_ONCE_INIT;

static void mk_init(void)
{
    fkm_mfma_kernel(16, &g_mk[0]);
    fkm_mfma_kernel(256, &g_mk[1]);
    fkm_mfma_kernel(4096, &g_mk[2]);
}

int fk_plan_use_mfma(fk_plan *p)
{
    fk_kernel *mk = g_mk;
    int i, j, n = 0;
    if (!p || p->prec != 1) return 0;
    pthread_once(&g_mk_once, mk_init);
    for (i = 0; i < p->npass; ++i) {
        fk_pass *ps = p->pass + i;
        const fk_kernel *k = ps->k;
        int all16 = 1;
        if (ps->type != FK_PASS_STOCKHAM || k->io == FK_IO_R2C || k->io == FK_IO_C2R ||
            k->io == FK_IO_BLUE || k->io == FK_IO_R2C_RR || k->io == FK_IO_C2R_RR ||
            k->io == FK_IO_R2C_XR || k->io == FK_IO_C2R_XR ||
            k->io == FK_IO_DCT4 || k->io == FK_IO_R2C_FULL || k->io == FK_IO_C2R_FULL)
            continue;
        for (j = 0; j < k->nrad; ++j) all16 &= k->rad[j] == 16;
        if (!all16 || (k->L != 16 && k->L != 256 && k->L != 4096)) continue;
        ps->k = &mk[k->L == 16 ? 0 : k->L == 256 ? 1 : 2];
        /* tiles of nl lines, tpl * nl threads, static LDS only */
        ps->block = (unsigned)(ps->k->tpl * ps->k->nl);
        ps->lds = 0;
        ps->grid = (unsigned)((ps->a.nlines + ps->k->nl - 1) / ps->k->nl);
        ps->a.xcd_remap = 1;
        ++n;
    }
    if (n) fk_count_ops(p);
    return n;
}

/* Enqueue every pass on `st` with the plan's own scratch; no host
   synchronisation (usable inside stream capture). */
int fk_enqueue(fk_plan *p, void *in, void *out, hipStream_t st)
{
    return fk_enqueue_bound(p, in, out, p->scratch, st);
}

/* Same with caller-owned scratch slots (scratch[i] >= p->scratch_bytes[i]):
   several executions of one plan may then overlap on the device. */
int fk_enqueue_bound(fk_plan *p, void *in, void *out, void *const *scratch, hipStream_t st)
{
    void *buf[FK_NBUF];
    int i, rc = 0;
    buf[FK_BUF_IN] = in;
    buf[FK_BUF_OUT] = out;
    buf[FK_BUF_IN2] = p->in2;
    buf[FK_BUF_OUT2] = p->out2;
    for (i = 0; i < FK_MAX_SCRATCH; ++i) buf[FK_BUF_S0 + i] = scratch[i];
    for (i = 0; i < p->npass && !rc; ++i) rc = launch_pass(p->pass + i, buf, st, p->prec);
    return rc;
}

void fk_execute_or_die(fk_plan *p, void *in, void *out)
{
    if (fk_execute(p, in, out)) {
        fprintf(stderr, "fftw3_amd: %s\n", fk_last_error());
        abort();
    }
}

/* Per-pass timing with HIP events on the plan's stream. */
double fk_plan_time(fk_plan *p, void *in, void *out, int iters, double *per_pass_ms)
{
    void *buf[FK_NBUF];
    hipEvent_t *ev;
    double total = 0;
    int i, it, rc = 0;
    hipStream_t st = p->stream;
    if (iters < 1) iters = 1;
    buf[FK_BUF_IN] = in;
    buf[FK_BUF_OUT] = out;
    buf[FK_BUF_IN2] = p->in2;
    buf[FK_BUF_OUT2] = p->out2;
    for (i = 0; i < FK_MAX_SCRATCH; ++i) buf[FK_BUF_S0 + i] = p->scratch[i];
    ev = (hipEvent_t *)calloc((size_t)p->npass + 1, sizeof(hipEvent_t));
    for (i = 0; i <= p->npass; ++i)
        if (hipEventCreate(&ev[i]) != hipSuccess) rc = -1;
    for (i = 0; i < p->npass && per_pass_ms; ++i) per_pass_ms[i] = 0;
    for (it = 0; it < iters && !rc; ++it) {
        float ms;
        rc = (int)hipEventRecord(ev[0], st);
        for (i = 0; i < p->npass && !rc; ++i) {
            rc = launch_pass(p->pass + i, buf, st, p->prec);
            if (!rc) rc = (int)hipEventRecord(ev[i + 1], st);
        }
        if (!rc) rc = (int)hipEventSynchronize(ev[p->npass]);
        for (i = 0; i < p->npass && !rc; ++i) {
            if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) != hipSuccess) rc = -1;
            if (per_pass_ms) per_pass_ms[i] += ms / iters;
        }
        if (!rc && hipEventElapsedTime(&ms, ev[0], ev[p->npass]) == hipSuccess) total += ms;
    }
    for (i = 0; i <= p->npass; ++i)
        if (ev[i]) (void)hipEventDestroy(ev[i]);
    free(ev);
    if (rc) {
        set_err("timing failed");
        return -1;
    }
    return total / iters;
}

/* Device ms of `iter` executes queued back to back on the planning arrays
   and the plan's stream, between two events (no host gap inside). */
static double fk_plan_time_run(fk_plan *p, int iter)
{
    void *buf[FK_NBUF];
    hipEvent_t e0 = 0, e1 = 0;
    hipStream_t st = p->stream;
    float ms = -1.f;
    int i, it, rc = 0;
    buf[FK_BUF_IN] = p->in;
    buf[FK_BUF_OUT] = p->out;
    buf[FK_BUF_IN2] = p->in2;
    buf[FK_BUF_OUT2] = p->out2;
    for (i = 0; i < FK_MAX_SCRATCH; ++i) buf[FK_BUF_S0 + i] = p->scratch[i];
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = -1;
    if (!rc) rc = (int)hipEventRecord(e0, st);
    for (it = 0; it < iter && !rc; ++it) {
        if (p->nrot > 1) {
            /* the rotation continues across calls (the planner lock is
               held), so no repeat starts on the copy the last one ended on */
            static unsigned pos;
            const int k = (int)(pos++ % (unsigned)p->nrot);
            buf[FK_BUF_IN] = p->rot[2 * k];
            buf[FK_BUF_OUT] = p->rot[2 * k + 1];
        }
        for (i = 0; i < p->npass && !rc; ++i) rc = launch_pass(p->pass + i, buf, st, p->prec);
    }
    if (!rc) rc = (int)hipEventRecord(e1, st);
    if (!rc) rc = (int)hipEventSynchronize(e1);
    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = -1;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (rc) {
        set_err("timing failed");
        return -1;
    }
    return ms;
}

/* ------------------------------------------------------------------ */
/* introspection                                                       */

int fk_plan_npasses(const fk_plan *p) { return p ? p->npass : 0; }

/* Modelled seconds of one execute, from the plan's passes: each pass takes
   the longer of its HBM bytes at the streaming rate and its arithmetic at
   the vector peak of its precision, plus a launch; a pass whose contiguous
   runs are short (strided columns of few lines) pays the penalty of
   seg_pen().  This is the planner-side estimate that fftw_estimate_cost
   reports (the role of kernel/planner.c:426-442's cost model). */
double fk_plan_estimate(const fk_plan *p)
{
    const double bw = 5.2e12, launch = 3e-6;
    const double peak = p && p->prec ? 157.3e12 : 78.6e12;
    double t = 0;
    int i;
    if (!p) return 0;
    for (i = 0; i < p->npass; ++i) {
        const fk_pass *ps = p->pass + i;
        double tb = ps->bytes / bw, tf = 0;
        if (ps->type == FK_PASS_STOCKHAM) {
            const fk_kernel *k = ps->k;
            double fl = 5.0 * k->L * log2((double)k->L) * (double)ps->a.nlines;
            if (k->mode == FK_COL) tb *= 1.0 + seg_pen((double)k->nl * csz(k->prec));
            tf = fl / peak;
        }
        t += (tb > tf ? tb : tf) + launch;
    }
    return t;
}

double fk_plan_bytes(const fk_plan *p)
{
    double b = 0;
    int i;
    for (i = 0; p && i < p->npass; ++i) b += p->pass[i].bytes;
    return b;
}

static const char *buf_name(int b)
{
    static const char *n[] = {"in", "out", "in2", "out2"};
    static const char *s[] = {"s0", "s1", "s2", "s3", "s4", "s5", "s6", "s7", "s8", "s9", "s10"};
    if (b >= 0 && b < FK_BUF_S0) return n[b];
    return b >= FK_BUF_S0 && b < FK_NBUF ? s[b - FK_BUF_S0] : "?";
}

static const char *io_name(int io)
{
    static const char *n[] = {"plain", "tout", "r2c", "c2r", "bluestein", "r2c-r2r", "c2r-r2r",
                              "dct4", "r2c-full", "c2r-full", "r2c-xr", "c2r-xr"};
    return io >= 0 && io < 12 ? n[io] : "?";
}

int fk_plan_pass_info(const fk_plan *p, int i, char *buf, int len)
{
    const fk_pass *ps;
    if (!p || i < 0 || i >= p->npass) return -1;
    ps = p->pass + i;
    if (ps->type == FK_PASS_STOCKHAM) {
        char rad[64] = "";
        int j;
        for (j = 0; j < ps->k->nrad; ++j)
            snprintf(rad + strlen(rad), sizeof(rad) - strlen(rad), "%s%d", j ? "x" : "",
                     ps->k->rad[j]);
        return snprintf(buf, len,
                        "%s %s L=%d rad=%s %s %s tpl=%d nl=%d%s lines=%lld grid=%u lds=%u "
                        "is=%lld os=%lld flags=%u src=%s dst=%s",
                        ps->k->variant >= 100 ? "mfma16" : "stockham", ps->k->prec ? "f32" : "f64", ps->k->L, rad,
                        ps->k->mode == FK_ROW ? "row" : "col", io_name(ps->k->io),
                        ps->k->tpl, ps->k->nl, ps->k->loop ? " pipelined" : "", ps->a.nlines, ps->grid, ps->lds, ps->a.is,
                        ps->a.os, ps->a.flags, buf_name(ps->ibuf), buf_name(ps->obuf));
    }
    return snprintf(buf, len, "elementwise op=%d n=%lld lines=%lld src=%s dst=%s", ps->e.op,
                    ps->e.n, ps->e.nlines, buf_name(ps->ibuf), buf_name(ps->obuf));
}

char *fk_plan_sprint(const fk_plan *p)
{
    size_t cap = 4096, len = 0;
    char *s = (char *)malloc(cap), line[512];
    int i;
    s[0] = 0;
    if (!p) return s;
    len = snprintf(s, cap, "(fftw3-amd-gfx950 %s %s\n", p->prec ? "float" : "double", p->desc);
    for (i = 0; i < p->npass; ++i) {
        size_t l;
        fk_plan_pass_info(p, i, line, sizeof(line));
        l = strlen(line) + 8;
        if (len + l >= cap) {
            cap = 2 * (len + l);
            s = (char *)realloc(s, cap);
        }
        len += snprintf(s + len, cap - len, "  (%s)\n", line);
    }
    if (len + 4 >= cap) s = (char *)realloc(s, len + 4);
    strcat(s, ")");
    return s;
}

void fk_plan_print(const fk_plan *p, FILE *f)
{
    char *s = fk_plan_sprint(p);
    fputs(s, f);
    free(s);
}

/* ------------------------------------------------------------------ */

void fk_cleanup(void)
{
    tbl *t, *n;
    pthread_mutex_lock(&g_lock);
    for (t = g_tbl; t; t = n) {
        n = t->next;
        fk_free(t->dev);
        free(t);
    }
    g_tbl = 0;
    pthread_mutex_unlock(&g_lock);
}
