/*
 * Project extensions of the MI355X engine.  These have no counterpart in
 * the reference's fftw3.h: they exist because the drop-in's arrays live in
 * device memory (copies, fills, streams) and for measurement.
 *
 * Plans of either precision (fftw_plan, fftwf_plan) are passed as void *.
 */
#ifndef FFTW3_AMD_H
#define FFTW3_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Initialise the device runtime and kernel catalog; 0 on success.  Called
   implicitly by every planner. */
int fftw_amd_init(void);

/* Last error message of this thread (planner returned NULL, execute failed). */
const char *fftw_amd_last_error(void);

/* Stream used by plans created from now on (hipStream_t, 0 = null stream),
   and the stream of one existing plan.  Switching a plan's stream makes the
   new stream wait for the work already queued on the old one, so async
   executes of one plan, which share its scratch, never overlap. */
void fftw_amd_set_stream(void *hip_stream);
void *fftw_amd_get_stream(void);
void fftw_amd_plan_set_stream(void *plan, void *hip_stream);

/* 0 (default): execute returns after the transform completed, as FFTW's
   execute does.  1: execute only enqueues work on the plan's stream; the
   plan's scratch is then reused by the next execute, which is safe because
   every execute of a plan goes to the plan's stream (in queue order). */
void fftw_amd_set_async(int on);

/* Device memory helpers; return 0 on success. */
int fftw_amd_memcpy_htod(void *dst, const void *src, size_t bytes);
int fftw_amd_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int fftw_amd_memcpy_dtod(void *dst, const void *src, size_t bytes);
int fftw_amd_memset(void *dst, int value, size_t bytes);
/* Streaming device copy (hand-written kernel: 16-byte accesses, four in
   flight per thread, `ctas` workgroups of 256 threads looping over the
   array, 0 = 2048); mode 0 plain, 1 non-temporal.  Enqueued on `stream`
   (hipStream_t, 0 = null stream), not synchronized.  16-byte aligned
   pointers; returns 0 or a HIP error code.  The benchmark's practical HBM
   ceiling. */
int fftw_amd_stream_copy(void *dst, const void *src, size_t bytes, int mode, int ctas,
                         void *hip_stream);
int fftw_amd_synchronize(void);
size_t fftw_amd_mem_free(void);

/* Counter-based synthetic inputs written on the device: element i of
   `count` reals is 2u-1 with u the top 53 bits of
   splitmix64(seed ^ (stream << 40) ^ i) / 2^53 (gauss: Box-Muller on
   elements 2i, 2i+1).  prec: 0 double, 1 float. */
int fftw_amd_fill_uniform(void *dev, int prec, size_t count, uint64_t seed, uint64_t stream);
int fftw_amd_fill_gauss(void *dev, int prec, size_t count, uint64_t seed, uint64_t stream);

/* Plan introspection */
int fftw_amd_plan_npasses(const void *plan);
int fftw_amd_plan_pass_info(const void *plan, int i, char *buf, int len);
double fftw_amd_plan_bytes(const void *plan);
/* Algorithmic HBM bytes of pass i (each input read once, output written once). */
double fftw_amd_plan_pass_bytes(const void *plan, int i);

/* Execute `plan` `iters` times on (in, out) with HIP events around every
   pass; writes the mean milliseconds of pass i to per_pass_ms[i] (array of
   fftw_amd_plan_npasses entries) and returns the mean milliseconds of a
   whole execute.  Negative on error. */
double fftw_amd_plan_time(void *plan, void *in, void *out, int iters, double *per_pass_ms);

/* Allocation-free execute.  No counterpart in the reference: it is for
   callers that cannot allocate device memory during the call (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   Runs `plan` on new arrays exactly as fftw_execute_dft / _r2c / _c2r /
   _r2r would, but every scratch buffer of the plan is carved out of one
   static device array of FFTW_AMD_STATIC_SCRATCH_BYTES (128 MiB, held by
   each device the library runs on until the process ends) instead of the
   plan's own scratch, and no allocation happens on any path.  Plans whose
   scratch exceeds that size, and split-array plans, are rejected with a
   nonzero HIP error code (fftw_amd_last_error says why).  All calls on one
   device share the array, which keeps the last call's contents: calls
   that overlap on one device (other host threads, other streams, or async
   mode with executes of different plans in flight on different streams)
   race.  Synchronizes unless fftw_amd_set_async(1).  Returns 0 on
   success. */
#define FFTW_AMD_STATIC_SCRATCH_BYTES ((size_t)128 << 20)
int fftw_amd_execute_noalloc(const void *plan, void *in, void *out);

/* Kernel catalog (the compiled Stockham kernels the planner chooses from).
   fftw_amd_catalog_entry fills info[22]: precision (0 double, 1 float),
   length L, mapping (0 row, 1 column), IO kind (0 plain, 1 transposed
   store, 2 r2c, 3 c2r, 4 Bluestein, 5 r2c with a real-to-real kind, 6 c2r
   with a real-to-real kind, 7 DCT-IV, 8 full-length r2c, 9 full-length
   c2r, 10 r2c with the first radix-(lines per workgroup) stage of the next
   dimension across the workgroup's rows, 11 c2r with the last inverse such
   stage), threads per line, lines per workgroup, variant (0 = the default
   geometry), number of radices, the radices (12 slots), dynamic LDS bytes,
   1 when each workgroup runs several tiles, loading the next tile's
   inputs during the current tile's stages (grid sized to the resident
   workgroups).
   Returns 0, or -1 for an index out of range. */
int fftw_amd_catalog_size(void);
int fftw_amd_catalog_entry(int i, int *info);
/* Test hook: plans created after this call use catalog entry i wherever
   the planner would pick a kernel of the same precision, length, mapping
   and IO kind (whichever variant it would have chosen); -1 restores normal
   planning.  Lets every catalog entry, including the variants only the
   autotuner picks, be run and checked.  Returns 0, or -1 for a bad index. */
int fftw_amd_force_kernel(int i);

/* Split-precision matrix-core variant (FP32 plans only): every pass of
   length 16, 256 or 4096 with radix-16 stages runs its butterflies as FP16
   hi/lo matrix products with FP32 accumulation (v_mfma_f32_32x32x8_f16).
   Returns the number of passes switched (0 for other plans). */
int fftw_amd_plan_use_mfma(void *fftwf_plan);
/* Planner flag for fftwf_ c2c plans: decompose power-of-16 lengths into
   16 / 256 / 4096-point passes and run them on the matrix-core variant. */
#define FFTW_AMD_MFMA (1U << 24)
/* Planner flag: plan multi-dimensional problems cache-blocked by default
   (passes run per chunk of transforms, or per slab of planes for unbatched
   rank >= 3, sized for the 256 MB memory-side cache), so intermediates are
   re-read from that cache rather than HBM.  FFTW_PATIENT also times the
   unblocked plan and keeps the faster. */
#define FFTW_AMD_BLOCKED (1U << 25)
/* Writes the 32x32 result of one probe matrix product (identity rows 0..7
   times a known B) to dev_out, for checking the operand layouts. */
int fftw_amd_mfma_probe(float *dev_out);

/* Batched transforms of uneven lengths.  Group g holds howmany[g]
   contiguous 1-D c2c transforms of length n[g] (distance n[g]) from in[g] to
   out[g] (device memory).  Every group is planned with `flags` (FFTW_MEASURE
   tunes each size on the device); the passes of all groups run as one
   persistent, work-table driven launch per phase, recorded once into a
   hipGraph and replayed by fftw_amd_batch_execute.  prec: 0 double, 1 float.
   Returns NULL on invalid arguments or allocation failure. */
typedef struct fftw_amd_batch_s *fftw_amd_batch;
fftw_amd_batch fftw_amd_plan_batch_dft_1d(int prec, int ngroups, const int *n, const int *howmany,
                                          void *const *in, void *const *out, int sign,
                                          unsigned flags);
int fftw_amd_batch_execute(fftw_amd_batch b);         /* graph replay */
int fftw_amd_batch_execute_direct(fftw_amd_batch b);  /* same launches, no graph */
int fftw_amd_batch_info(fftw_amd_batch b, char *buf, int len);
void *fftw_amd_batch_stream(fftw_amd_batch b);
void fftw_amd_batch_destroy(fftw_amd_batch b);

/* Serving loop with continuous batching.
   A request: kind (0 c2c, 1 r2c, 2 c2r), prec (0 double, 1 float), sign
   (c2c), rank 1..3 with sizes n[], howmany contiguous row-major transforms
   (complex side of r2c/c2r has n/2+1 along the last dim), input in device
   memory (unchanged until the request completes; multi-dimensional c2r
   destroys it, as in FFTW).  The server owns the outputs: an arena of
   `hbm_budget` bytes holds every admitted request's output and scratch;
   requests are admitted in arrival order as budget and round slots free
   up; up to `max_rounds` rounds are in flight, each with up to
   `per_round` requests, whose menu-compatible passes run as one persistent
   launch per phase and precision. */
typedef struct {
    int kind, prec, sign, rank;
    int n[3];
    int howmany;
    const void *in;
} fftw_amd_request;
typedef struct fftw_amd_server_s *fftw_amd_server;
fftw_amd_server fftw_amd_server_create(size_t hbm_budget, int max_rounds, int per_round);
/* plan a request shape ahead of time (plans stay resident); 0 on success */
int fftw_amd_server_prepare(fftw_amd_server s, const fftw_amd_request *r);
/* request id >= 0; -1 if invalid or larger than the whole budget */
long long fftw_amd_server_submit(fftw_amd_server s, const fftw_amd_request *r);
int fftw_amd_server_poll(fftw_amd_server s, long long id);    /* 1 done, 0 not yet, -1 error */
/* block until done; *out = device pointer of the result, *latency_ms =
   completion - submission */
int fftw_amd_server_wait(fftw_amd_server s, long long id, void **out, double *latency_ms);
int fftw_amd_server_release(fftw_amd_server s, long long id); /* returns the output's bytes */
int fftw_amd_server_stats(fftw_amd_server s, double *p50_ms, double *p99_ms, long long *completed,
                          long long *transforms, double *elapsed_s, size_t *hbm_in_use);
void fftw_amd_server_destroy(fftw_amd_server s);

/* Events on the current stream for timing regions (milliseconds). */
void *fftw_amd_event_create(void);
void fftw_amd_event_record(void *ev, void *hip_stream);
double fftw_amd_event_elapsed(void *start, void *stop);
void fftw_amd_event_destroy(void *ev);

#ifdef __cplusplus
}
#endif
#endif
