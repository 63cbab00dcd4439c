/*
 * CPU oracle for the DFT hot path.  TEST INFRASTRUCTURE ONLY: nothing in the
 * product library links this; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it, and only to check results.
 *
 * It restates, from the mathematics, what FFTW 3.3 computes on the
 * plan/execute path:
 *   - the DFT definition: Y_k = sum_j X_j exp(sign * 2*pi*i*j*k/n),
 *     unnormalized, sign = -1 forward / +1 backward
 *     (doc/reference.texi:1872-1902, kernel/ifftw.h:92 FFT_SIGN=-1);
 *   - guru tensor semantics: rank dims {n, is, os} plus howmany dims
 *     {n, is, os}; element at sum(index_d * stride_d)
 *     (api/mktensor-iodims.h:23-38, doc/reference.texi:1062-1070);
 *   - r2c: n reals -> n_last/2+1 complex along the last dim; c2r: the
 *     unnormalized inverse, computed as the backward complex DFT over the
 *     leading dims followed by a 1-D c2r along the last dim, where the 1-D
 *     c2r reads only Re(X[0]) and, for even n, Re(X[n/2])
 *     (rdft/rank-geq2-rdft2.c:55-69, doc/tutorial.texi:296-366);
 *   - twiddles exp(2*pi*i*m/n) computed with exact integer reduction of m
 *     and octant symmetry before cos/sin (kernel/trig.c:57-80);
 *   - Bluestein chirp w_m = exp(i*pi*m^2/n) with m^2 reduced mod 2n
 *     (dft/bluestein.c:37-50);
 *   - r2r kinds: halfcomplex R2HC / HC2R, DHT, DCT-I..IV (REDFTab) and
 *     DST-I..IV (RODFTab) as defined in doc/reference.texi:913-947 and
 *     :2021-2350, each computed as a complex DFT of its logical symmetric
 *     sequence (length 2(n-1), 2(n+1), 4n or 8n).
 *
 * Internal precision: ORC_LD computes in x87 long double (64-bit mantissa),
 * the accuracy oracle; ORC_D computes the same algorithm in double, which is
 * what the CPU throughput baseline times.  All arrays crossing this API are
 * double (float data is widened exactly by the caller).
 */
#ifndef FFTW_AMD_ORACLE_H
#define FFTW_AMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_LD = 0, ORC_D = 1 };

/* Strides are in elements of the array they index: complex elements for
   complex arrays, reals for real arrays. */
typedef struct { long n, is, os; } orc_iodim;

/* Complex DFT, guru form.  in/out: interleaved (re, im) doubles.  in == out
   is allowed (any strides).  Returns 0 on success, -1 on bad arguments. */
int orc_dft(int prec, int rank, const orc_iodim *dims, int hrank,
            const orc_iodim *hdims, const double *in, double *out, int sign);

/* Real-to-complex (forward).  in: reals; out: complex, last dim n/2+1. */
int orc_r2c(int prec, int rank, const orc_iodim *dims, int hrank,
            const orc_iodim *hdims, const double *in, double *out);

/* Complex-to-real (backward, unnormalized).  in: complex, last dim n/2+1;
   out: reals.  The input is not modified. */
int orc_c2r(int prec, int rank, const orc_iodim *dims, int hrank,
            const orc_iodim *hdims, const double *in, double *out);

/* Real-to-real transforms, guru form, kinds[d] = fftw_r2r_kind value of
   dim d (0 R2HC, 1 HC2R, 2 DHT, 3-6 REDFT00/01/10/11, 7-10 RODFT00/01/10/11).
   Reals in and out; the input is not modified.  -1 on bad arguments,
   including a size-1 REDFT00 dim. */
int orc_r2r(int prec, int rank, const orc_iodim *dims, int hrank,
            const orc_iodim *hdims, const int *kinds, const double *in,
            double *out);

/* Brute-force O(n^2) DFT in long double with exact (j*k mod n) reduction.
   Contiguous interleaved complex in/out (out != in). */
void orc_dft_direct(long n, int sign, const double *in, double *out);

/* Direct sums for selected output bins of a length-n contiguous DFT, long
   double with exact index reduction: out[2*t..] = Y[bins[t]]. */
void orc_dft_bins(long n, int sign, const double *in, const long *bins,
                  long nbins, double *out);

/* Twiddle exp(sign*2*pi*i*m/n) in long double, octant-reduced; written as
   two doubles (the rounded value). */
void orc_cexp(long m, long n, int sign, double *out2);
/* sampled bins (rank-tuples in `bins`) of the rank-r DFT of a row-major
   complex (real = 0) or real (real = 1) array, long double sums */
void orc_dft_bins_nd(int rank, const long *n, int sign, const double *in, int real,
                     const long *bins, long nbins, double *out);

/* Counter-based inputs: x[i] = 2*u - 1, u = top 53 bits of
   splitmix64(seed ^ (stream << 40) ^ i) / 2^53.  Same stream on the GPU. */
uint64_t orc_splitmix64(uint64_t key);
void orc_fill_uniform(double *x, long count, uint64_t seed, uint64_t stream);
/* Gaussian N(0,1): Box-Muller on uniforms of indices 2i and 2i+1. */
void orc_fill_gauss(double *x, long count, uint64_t seed, uint64_t stream);

/* Number of threads the oracle uses for batched work (OpenMP). */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
