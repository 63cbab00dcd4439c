/*
 * fftw3.h for the MI355X (gfx950) engine.
 *
 * Drop-in for FFTW 3.3's public plan/execute interface: same function
 * names, signatures, types and flag values as the reference header
 * (api/fftw3.h:127-463 declarations, :468-469 double/float instances,
 * :489-519 constants).  The difference in contract: every array argument
 * is DEVICE memory (hipMalloc / fftw_malloc), and the transform runs on the
 * GPU.  fftw_malloc returns device memory.
 *
 * Per-function reference citations (api/fftw3.h line of the declaration):
 *   execute :142          plan_dft :145          plan_dft_1d/2d/3d :149-157
 *   plan_many_dft :159    plan_guru_dft :168     plan_guru_split_dft :173
 *   plan_guru64_dft :181  execute_dft :196       execute_split_dft :199
 *   plan_many_dft_r2c :203 plan_dft_r2c{,_1d,_2d,_3d} :212-226
 *   plan_many_dft_c2r :228 plan_dft_c2r{,_1d,_2d,_3d} :237-251
 *   plan_guru_dft_r2c/c2r :253-264 plan_guru64_dft_r2c/c2r :279-292
 *   execute_dft_r2c/c2r :310-314 copy_plan :369 destroy_plan :372
 *   forget_wisdom :375 cleanup :377 set_timelimit :380
 *   threads api :382-397 wisdom export/import :399-427
 *   print_plan :429-436 malloc/alloc/free :438-447
 *   flops/estimate_cost/cost :449-456 alignment_of :458 version :461
 *
 * Split-complex guru planners (plan_guru{,64}_split_dft{,_r2c,_c2r}) and
 * execute_split_dft{,_r2c,_c2r} follow api/plan-guru-split-dft*.c and
 * api/execute-split-dft*.c: strides count reals, no sign argument (swap the
 * real and imaginary pointers of each pair for the backward transform).
 *
 * Real-to-real planners (plan_r2r{,_1d,_2d,_3d}, plan_many_r2r,
 * plan_guru{,64}_r2r, execute_r2r) follow api/plan-many-r2r.c:26-51,
 * api/plan-guru-r2r.h:23-44 and api/execute-r2r.c:24-28: every kind of
 * api/fftw3.h:89-100 (R2HC, HC2R, DHT, REDFT00-11, RODFT00-11), any rank,
 * strides in reals.  A size-1 REDFT00 dim returns NULL
 * (doc/reference.texi:818-821).  The input is preserved.
 */
#ifndef FFTW3_H
#define FFTW3_H

#include <stdio.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#if !defined(FFTW_NO_Complex) && defined(_Complex_I) && defined(complex) && defined(I)
#define FFTW_AMD_COMPLEX(R, C) typedef R _Complex C
#else
#define FFTW_AMD_COMPLEX(R, C) typedef R C[2]
#endif

enum fftw_r2r_kind_do_not_use_me {
    FFTW_R2HC = 0, FFTW_HC2R = 1, FFTW_DHT = 2,
    FFTW_REDFT00 = 3, FFTW_REDFT01 = 4, FFTW_REDFT10 = 5, FFTW_REDFT11 = 6,
    FFTW_RODFT00 = 7, FFTW_RODFT01 = 8, FFTW_RODFT10 = 9, FFTW_RODFT11 = 10
};

struct fftw_iodim_do_not_use_me { int n; int is; int os; };
struct fftw_iodim64_do_not_use_me { ptrdiff_t n; ptrdiff_t is; ptrdiff_t os; };

typedef void (*fftw_write_char_func_do_not_use_me)(char c, void *);
typedef int (*fftw_read_char_func_do_not_use_me)(void *);

/* One precision instance: P = prefix token, R = real, C = complex. */
#define FFTW_AMD_API(P, R, C)                                                               \
    FFTW_AMD_COMPLEX(R, C);                                                                 \
    typedef struct P##plan_s *P##plan;                                                      \
    typedef struct fftw_iodim_do_not_use_me P##iodim;                                       \
    typedef struct fftw_iodim64_do_not_use_me P##iodim64;                                   \
    typedef enum fftw_r2r_kind_do_not_use_me P##r2r_kind;                                   \
    typedef fftw_write_char_func_do_not_use_me P##write_char_func;                          \
    typedef fftw_read_char_func_do_not_use_me P##read_char_func;                            \
    extern void P##execute(const P##plan p);                                                \
    extern P##plan P##plan_dft(int rank, const int *n, C *in, C *out, int sign,             \
                               unsigned flags);                                             \
    extern P##plan P##plan_dft_1d(int n, C *in, C *out, int sign, unsigned flags);          \
    extern P##plan P##plan_dft_2d(int n0, int n1, C *in, C *out, int sign, unsigned flags); \
    extern P##plan P##plan_dft_3d(int n0, int n1, int n2, C *in, C *out, int sign,          \
                                  unsigned flags);                                          \
    extern P##plan P##plan_many_dft(int rank, const int *n, int howmany, C *in,             \
                                    const int *inembed, int istride, int idist, C *out,     \
                                    const int *onembed, int ostride, int odist, int sign,   \
                                    unsigned flags);                                        \
    extern P##plan P##plan_guru_dft(int rank, const P##iodim *dims, int howmany_rank,       \
                                    const P##iodim *howmany_dims, C *in, C *out, int sign,  \
                                    unsigned flags);                                        \
    extern P##plan P##plan_guru_split_dft(int rank, const P##iodim *dims, int howmany_rank, \
                                          const P##iodim *howmany_dims, R *ri, R *ii,       \
                                          R *ro, R *io, unsigned flags);                    \
    extern P##plan P##plan_guru64_dft(int rank, const P##iodim64 *dims, int howmany_rank,   \
                                      const P##iodim64 *howmany_dims, C *in, C *out,        \
                                      int sign, unsigned flags);                            \
    extern P##plan P##plan_guru64_split_dft(int rank, const P##iodim64 *dims,               \
                                            int howmany_rank,                               \
                                            const P##iodim64 *howmany_dims, R *ri, R *ii,   \
                                            R *ro, R *io, unsigned flags);                  \
    extern void P##execute_dft(const P##plan p, C *in, C *out);                             \
    extern void P##execute_split_dft(const P##plan p, R *ri, R *ii, R *ro, R *io);          \
    extern P##plan P##plan_many_dft_r2c(int rank, const int *n, int howmany, R *in,         \
                                        const int *inembed, int istride, int idist, C *out, \
                                        const int *onembed, int ostride, int odist,         \
                                        unsigned flags);                                    \
    extern P##plan P##plan_dft_r2c(int rank, const int *n, R *in, C *out, unsigned flags);  \
    extern P##plan P##plan_dft_r2c_1d(int n, R *in, C *out, unsigned flags);                \
    extern P##plan P##plan_dft_r2c_2d(int n0, int n1, R *in, C *out, unsigned flags);       \
    extern P##plan P##plan_dft_r2c_3d(int n0, int n1, int n2, R *in, C *out,                \
                                      unsigned flags);                                      \
    extern P##plan P##plan_many_dft_c2r(int rank, const int *n, int howmany, C *in,         \
                                        const int *inembed, int istride, int idist, R *out, \
                                        const int *onembed, int ostride, int odist,         \
                                        unsigned flags);                                    \
    extern P##plan P##plan_dft_c2r(int rank, const int *n, C *in, R *out, unsigned flags);  \
    extern P##plan P##plan_dft_c2r_1d(int n, C *in, R *out, unsigned flags);                \
    extern P##plan P##plan_dft_c2r_2d(int n0, int n1, C *in, R *out, unsigned flags);       \
    extern P##plan P##plan_dft_c2r_3d(int n0, int n1, int n2, C *in, R *out,                \
                                      unsigned flags);                                      \
    extern P##plan P##plan_guru_dft_r2c(int rank, const P##iodim *dims, int howmany_rank,   \
                                        const P##iodim *howmany_dims, R *in, C *out,        \
                                        unsigned flags);                                    \
    extern P##plan P##plan_guru_dft_c2r(int rank, const P##iodim *dims, int howmany_rank,   \
                                        const P##iodim *howmany_dims, C *in, R *out,        \
                                        unsigned flags);                                    \
    extern P##plan P##plan_guru_split_dft_r2c(int rank, const P##iodim *dims,               \
                                              int howmany_rank,                             \
                                              const P##iodim *howmany_dims, R *in, R *ro,   \
                                              R *io, unsigned flags);                       \
    extern P##plan P##plan_guru_split_dft_c2r(int rank, const P##iodim *dims,               \
                                              int howmany_rank,                             \
                                              const P##iodim *howmany_dims, R *ri, R *ii,   \
                                              R *out, unsigned flags);                      \
    extern P##plan P##plan_guru64_dft_r2c(int rank, const P##iodim64 *dims,                 \
                                          int howmany_rank,                                 \
                                          const P##iodim64 *howmany_dims, R *in, C *out,    \
                                          unsigned flags);                                  \
    extern P##plan P##plan_guru64_dft_c2r(int rank, const P##iodim64 *dims,                 \
                                          int howmany_rank,                                 \
                                          const P##iodim64 *howmany_dims, C *in, R *out,    \
                                          unsigned flags);                                  \
    extern P##plan P##plan_guru64_split_dft_r2c(int rank, const P##iodim64 *dims,           \
                                                int howmany_rank,                           \
                                                const P##iodim64 *howmany_dims, R *in,      \
                                                R *ro, R *io, unsigned flags);              \
    extern P##plan P##plan_guru64_split_dft_c2r(int rank, const P##iodim64 *dims,           \
                                                int howmany_rank,                           \
                                                const P##iodim64 *howmany_dims, R *ri,      \
                                                R *ii, R *out, unsigned flags);             \
    extern void P##execute_dft_r2c(const P##plan p, R *in, C *out);                         \
    extern void P##execute_dft_c2r(const P##plan p, C *in, R *out);                         \
    extern void P##execute_split_dft_r2c(const P##plan p, R *in, R *ro, R *io);             \
    extern void P##execute_split_dft_c2r(const P##plan p, R *ri, R *ii, R *out);            \
    extern P##plan P##plan_many_r2r(int rank, const int *n, int howmany, R *in,             \
                                    const int *inembed, int istride, int idist, R *out,     \
                                    const int *onembed, int ostride, int odist,             \
                                    const P##r2r_kind *kind, unsigned flags);               \
    extern P##plan P##plan_r2r(int rank, const int *n, R *in, R *out,                       \
                               const P##r2r_kind *kind, unsigned flags);                    \
    extern P##plan P##plan_r2r_1d(int n, R *in, R *out, P##r2r_kind kind, unsigned flags);  \
    extern P##plan P##plan_r2r_2d(int n0, int n1, R *in, R *out, P##r2r_kind kind0,         \
                                  P##r2r_kind kind1, unsigned flags);                       \
    extern P##plan P##plan_r2r_3d(int n0, int n1, int n2, R *in, R *out,                    \
                                  P##r2r_kind kind0, P##r2r_kind kind1, P##r2r_kind kind2,  \
                                  unsigned flags);                                          \
    extern P##plan P##plan_guru_r2r(int rank, const P##iodim *dims, int howmany_rank,       \
                                    const P##iodim *howmany_dims, R *in, R *out,            \
                                    const P##r2r_kind *kind, unsigned flags);               \
    extern P##plan P##plan_guru64_r2r(int rank, const P##iodim64 *dims, int howmany_rank,   \
                                      const P##iodim64 *howmany_dims, R *in, R *out,        \
                                      const P##r2r_kind *kind, unsigned flags);             \
    extern void P##execute_r2r(const P##plan p, R *in, R *out);                             \
    extern P##plan P##copy_plan(P##plan p);                                                 \
    extern void P##destroy_plan(P##plan p);                                                 \
    extern void P##forget_wisdom(void);                                                     \
    extern void P##cleanup(void);                                                           \
    extern void P##set_timelimit(double t);                                                 \
    extern void P##plan_with_nthreads(int nthreads);                                        \
    extern int P##planner_nthreads(void);                                                   \
    extern int P##init_threads(void);                                                       \
    extern void P##cleanup_threads(void);                                                   \
    extern void P##threads_set_callback(                                                    \
        void (*parallel_loop)(void *(*work)(char *), char *jobdata, size_t elsize,          \
                              int njobs, void *data),                                       \
        void *data);                                                                        \
    extern void P##make_planner_thread_safe(void);                                          \
    extern int P##export_wisdom_to_filename(const char *filename);                          \
    extern void P##export_wisdom_to_file(FILE *output_file);                                \
    extern char *P##export_wisdom_to_string(void);                                          \
    extern void P##export_wisdom(P##write_char_func write_char, void *data);                \
    extern int P##import_system_wisdom(void);                                               \
    extern int P##import_wisdom_from_filename(const char *filename);                        \
    extern int P##import_wisdom_from_file(FILE *input_file);                                \
    extern int P##import_wisdom_from_string(const char *input_string);                      \
    extern int P##import_wisdom(P##read_char_func read_char, void *data);                   \
    extern void P##fprint_plan(const P##plan p, FILE *output_file);                         \
    extern void P##print_plan(const P##plan p);                                             \
    extern char *P##sprint_plan(const P##plan p);                                           \
    extern void *P##malloc(size_t n);                                                       \
    extern R *P##alloc_real(size_t n);                                                      \
    extern C *P##alloc_complex(size_t n);                                                   \
    extern void P##free(void *p);                                                           \
    extern void P##flops(const P##plan p, double *add, double *mul, double *fmas);          \
    extern double P##estimate_cost(const P##plan p);                                        \
    extern double P##cost(const P##plan p);                                                 \
    extern int P##alignment_of(R *p);                                                       \
    extern const char P##version[];                                                         \
    extern const char P##cc[];                                                              \
    extern const char P##codelet_optim[];

FFTW_AMD_API(fftw_, double, fftw_complex)
FFTW_AMD_API(fftwf_, float, fftwf_complex)

#define FFTW_FORWARD (-1)
#define FFTW_BACKWARD (+1)

#define FFTW_NO_TIMELIMIT (-1.0)

/* documented flags (api/fftw3.h:495-504) */
#define FFTW_MEASURE (0U)
#define FFTW_DESTROY_INPUT (1U << 0)
#define FFTW_UNALIGNED (1U << 1)
#define FFTW_CONSERVE_MEMORY (1U << 2)
#define FFTW_EXHAUSTIVE (1U << 3)
#define FFTW_PRESERVE_INPUT (1U << 4)
#define FFTW_PATIENT (1U << 5)
#define FFTW_ESTIMATE (1U << 6)
#define FFTW_WISDOM_ONLY (1U << 21)

/* beyond-guru flags (api/fftw3.h:506-519); accepted and ignored */
#define FFTW_ESTIMATE_PATIENT (1U << 7)
#define FFTW_BELIEVE_PCOST (1U << 8)
#define FFTW_NO_DFT_R2HC (1U << 9)
#define FFTW_NO_NONTHREADED (1U << 10)
#define FFTW_NO_BUFFERING (1U << 11)
#define FFTW_NO_INDIRECT_OP (1U << 12)
#define FFTW_ALLOW_LARGE_GENERIC (1U << 13)
#define FFTW_NO_RANK_SPLITS (1U << 14)
#define FFTW_NO_VRANK_SPLITS (1U << 15)
#define FFTW_NO_VRECURSE (1U << 16)
#define FFTW_NO_SIMD (1U << 17)
#define FFTW_NO_SLOW (1U << 18)
#define FFTW_NO_FIXED_RADIX_LARGE_N (1U << 19)
#define FFTW_ALLOW_PRUNING (1U << 20)

#ifdef __cplusplus
}
#endif
#endif /* FFTW3_H */
