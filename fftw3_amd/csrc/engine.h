/*
 * Host engine shared by the double (fftw_) and float (fftwf_) API
 * instances.  A plan is a list of passes (kernel launches) over symbolic
 * buffers {input, output, scratch slots}; execute binds the user's device
 * pointers and launches the list on a HIP stream.
 *
 * Reference counterparts: problem/tensor types (dft/dft.h:34-38,
 * rdft/rdft.h:117-124, kernel/ifftw.h:417-432), the apiplan
 * (api/api.h:62-67) and the planner (kernel/planner.c:623-747), which here
 * becomes a pass-list builder plus a GPU timing autotuner.
 */
#ifndef FK_ENGINE_H
#define FK_ENGINE_H

#include <stddef.h>
#include <stdio.h>
#include <hip/hip_runtime_api.h>
#include "fk_args.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FK_MAXRANK 16
#define FK_MAX_SCRATCH 11
#define FK_SPLIT_SLOT 8         /* interleaved staging of split-array plans */
#define FK_R2R_SLOT 9           /* r2r staging: slot 9 pre-pass output, 10 inner output */

enum { FK_C2C = 0, FK_R2C = 1, FK_C2R = 2, FK_R2R = 3 };
/* r2r kinds per dimension, same values as fftw3.h's fftw_r2r_kind
   (api/fftw3.h:89-100) */
enum { FK_K_R2HC = 0, FK_K_HC2R = 1, FK_K_DHT = 2,
       FK_K_REDFT00 = 3, FK_K_REDFT01 = 4, FK_K_REDFT10 = 5, FK_K_REDFT11 = 6,
       FK_K_RODFT00 = 7, FK_K_RODFT01 = 8, FK_K_RODFT10 = 9, FK_K_RODFT11 = 10 };
/* symbolic buffers of a pass: the arrays given to execute (IN2 / OUT2 are
   the imaginary arrays of split-complex plans) and the plan's scratch */
enum { FK_BUF_IN = 0, FK_BUF_OUT = 1, FK_BUF_IN2 = 2, FK_BUF_OUT2 = 3, FK_BUF_S0 = 4 };
#define FK_NBUF (FK_BUF_S0 + FK_MAX_SCRATCH)

/* split-complex modes of a plan (api/plan-guru-split-dft.c: no sign
   argument; the transform is forward on (ri, ii) -> (ro, io)) */
enum { FK_SPLIT_NONE = 0,     /* interleaved arrays */
       FK_SPLIT_PAIRED = 1,   /* ii = ri + 1, io = ro + 1: run as interleaved */
       FK_SPLIT_SWAPPED = 2,  /* ri = ii + 1, ro = io + 1: interleaved backward */
       FK_SPLIT_GENERAL = 3 };/* separate arrays: gather / transform / scatter */
enum { FK_PASS_STOCKHAM = 0, FK_PASS_EW = 1 };

/* planner flags, same values as fftw3.h (api/fftw3.h:495-519) */
#define FK_MEASURE (0U)
#define FK_DESTROY_INPUT (1U << 0)
#define FK_UNALIGNED (1U << 1)
#define FK_CONSERVE_MEMORY (1U << 2)
#define FK_EXHAUSTIVE (1U << 3)
#define FK_PRESERVE_INPUT (1U << 4)
#define FK_PATIENT (1U << 5)
#define FK_ESTIMATE (1U << 6)
#define FK_WISDOM_ONLY (1U << 21)
#define FK_MFMA (1U << 24)          /* project flag: split-precision radix-16 */
#define FK_BLOCKED (1U << 25)       /* project flag: cache-blocked multi-D plans */

typedef struct { long long n, is, os; } fk_dim;

typedef struct {
    int type;
    const fk_kernel *k;
    fk_args a;
    fk_ew_args e;
    int ibuf, obuf;
    long long ioff, ooff;       /* byte offsets from the bound buffers */
    int i2buf, o2buf;           /* second input / output (split ops), -1 none */
    long long i2off, o2off;
    unsigned grid, block, lds;
    double bytes;               /* HBM bytes the pass moves (estimate) */
} fk_pass;

typedef struct fk_plan {
    int kind, prec, sign;
    unsigned flags;
    int rank, vrank;
    fk_dim dims[FK_MAXRANK], vdims[FK_MAXRANK];
    int inplace;
    void *in, *out;             /* arrays given at planning */
    int split;                  /* FK_SPLIT_* */
    void *in2, *out2;           /* imaginary arrays at planning (FK_SPLIT_GENERAL) */
    int r2r[FK_MAXRANK];        /* FK_R2R: kind of each dim */
    int npass, cpass;
    fk_pass *pass;
    int nscratch;
    size_t scratch_bytes[FK_MAX_SCRATCH];
    void *scratch[FK_MAX_SCRATCH];
    int scratch_busy;
    unsigned align_need;        /* required alignment of in/out (bytes) */
    unsigned align_in, align_out; /* more for one array, where a pass reads or
                                     writes it in pairs or quads */
    hipStream_t stream;
    int refcount;
    double adds, muls, fmas;    /* operation counts of the passes (opcount.c) */
    double pcost;               /* measured seconds per execute (MEASURE) */
    char desc[2048];            /* plan description for print_plan */
    int choice[32];             /* planner decisions (wisdom payload) */
    int nchoice;
    int tuned;                  /* 0 heuristic, 1 from wisdom, 2 timed now */
    /* while FFTW_MEASURE times the plan: nrot > 1 copies of its arrays,
       rot[2i] / rot[2i + 1] the input / output of copy i, used in turn */
    void *const *rot;
    int nrot;
} fk_plan;

/* problem description handed to the planner (strides in elements of the
   array they index, FFTW tensor semantics) */
typedef struct {
    int kind, prec, sign;
    unsigned flags;
    int rank, vrank;
    fk_dim dims[FK_MAXRANK], vdims[FK_MAXRANK];
    void *in, *out;
    /* split-complex arrays (FK_SPLIT_GENERAL): strides of every dim count
       reals; in2 / out2 are the imaginary arrays of the complex side(s) */
    int split;
    void *in2, *out2;
    /* FK_R2R: kind (FK_K_*) of each transform dim; strides count reals */
    int r2r[FK_MAXRANK];
} fk_problem;

/* ---- engine entry points ---- */
int fk_init(void);
fk_plan *fk_plan_create(const fk_problem *pb);
void fk_plan_destroy(fk_plan *p);
int fk_execute(fk_plan *p, void *in, void *out);         /* 0 or hip error */
/* split-complex arrays of an FK_SPLIT_GENERAL plan (in2 / out2 unused
   otherwise) */
int fk_execute4(fk_plan *p, void *in, void *in2, void *out, void *out2);
void fk_execute_or_die(fk_plan *p, void *in, void *out);
/* Execute with the static device scratch of fk_noalloc.hip in place of the
   plan's own: allocates nothing on any path.  Calls that overlap on one
   device share that scratch and race. */
#define FK_STATIC_SCRATCH_BYTES ((size_t)128 << 20)
int fk_execute_noalloc(fk_plan *p, void *in, void *out);
int fk_static_scratch_setup(void);
int fk_static_scratch(void **ptr, size_t *bytes);
void fk_cleanup(void);
void fk_forget_wisdom(void);
void fk_set_timelimit(double t);

/* device memory */
void *fk_malloc(size_t n);
void fk_free(void *p);

/* wisdom (text) */
char *fk_wisdom_export(void);                 /* malloc'd string */
int fk_wisdom_import(const char *s);          /* 1 ok, 0 rejected */
void fk_wisdom_store(unsigned long long h, int level, const int *c, int n);
int fk_wisdom_lookup(unsigned long long h, int *level, int *c, int *n);
int fk_wisdom_count(void);
unsigned long long fk_fnv1a(const void *p, size_t n, unsigned long long h);

/* extensions */
void fk_set_stream(hipStream_t s);
hipStream_t fk_get_stream(void);
void fk_set_async(int on);
int fk_get_async(void);
const char *fk_last_error(void);
void fk_plan_set_stream(fk_plan *p, hipStream_t s);
void fk_plan_print(const fk_plan *p, FILE *f);
char *fk_plan_sprint(const fk_plan *p);
int fk_plan_npasses(const fk_plan *p);
int fk_plan_pass_info(const fk_plan *p, int i, char *buf, int len);
double fk_plan_bytes(const fk_plan *p);
double fk_plan_estimate(const fk_plan *p);
/* kernel catalog introspection and the planner's forced-entry test hook */
int fk_catalog_size(void);
const fk_kernel *fk_catalog_entry(int i);
int fk_force_kernel(int i);
/* operation counts of the plan's passes (opcount.c) */
void fk_count_ops(fk_plan *p);
void fk_dft_ops(int n, double *add, double *mul, double *fma);
double fk_plan_time(fk_plan *p, void *in, void *out, int iters, double *per_pass_ms);

/* batched uneven sizes (batch.c) */
typedef struct fk_batch fk_batch;
fk_batch *fk_batch_create(int prec, int ngroups, const int *n, const int *howmany,
                          void *const *in, void *const *out, int sign, unsigned flags);
int fk_batch_execute(fk_batch *b, int use_graph);
int fk_batch_info(fk_batch *b, char *buf, int len);
hipStream_t fk_batch_stream(fk_batch *b);
void fk_batch_destroy(fk_batch *b);
int fk_plan_use_mfma(fk_plan *p);
int fkm_mfma_probe(float *dev_out, hipStream_t st);
int fk_enqueue(fk_plan *p, void *in, void *out, hipStream_t st);
int fk_enqueue_bound(fk_plan *p, void *in, void *out, void *const *scratch, hipStream_t st);
int fkm_reset_counters(unsigned long long *c, int n, hipStream_t st);

/* serving loop (serve.c) */
typedef struct {
    int kind, prec, sign, rank;
    int n[3];
    int howmany;
    const void *in;
} fk_request;
typedef struct fk_server fk_server;
fk_server *fk_server_create(size_t budget, int max_rounds, int per_round);
int fk_server_prepare(fk_server *s, const fk_request *r);
long long fk_server_submit(fk_server *s, const fk_request *r);
int fk_server_poll(fk_server *s, long long id);
int fk_server_wait(fk_server *s, long long id, void **out, double *lat_ms);
int fk_server_release(fk_server *s, long long id);
int fk_server_stats(fk_server *s, double *p50, double *p99, long long *done, long long *transforms,
                    double *elapsed_s, size_t *hbm_in_use);
void fk_server_destroy(fk_server *s);

/* launchers from fk_misc.hip */
int fkm_gen_stage_tw(void *out, int prec, int nrad, const int *rad, int total, hipStream_t st);
int fkm_gen_tw2(double *lo, double *hi, long long N, long long b, hipStream_t st);
int fkm_gen_line_tw(void *out, int prec, long long n, long long count, hipStream_t st);
int fkm_gen_chirp(void *out, int prec, long long n, hipStream_t st);
int fkm_gen_blue_b(void *out, long long n, long long M, hipStream_t st);
int fkm_stream_copy(void *dst, const void *src, long long bytes, int mode, int ctas, hipStream_t st);
int fkm_scale_round(void *out, int prec, const void *in, long long count, double scale,
                    hipStream_t st);
int fkm_fill(void *out, int prec, long long count, unsigned long long seed,
             unsigned long long stream, int gauss, hipStream_t st);
int fkm_ew(int prec, const fk_ew_args *a, hipStream_t st);

#ifdef __cplusplus
}
#endif
#endif
