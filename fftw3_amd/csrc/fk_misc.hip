/*
 * Support kernels and their C-ABI launchers:
 *   - twiddle, chirp and Bluestein-kernel tables, generated in FP64 on the
 *     GPU and rounded to the working precision (the role of
 *     kernel/twiddle.c:124-256 and dft/bluestein.c:37-80);
 *   - elementwise passes used between Stockham passes on the multi-pass,
 *     Bluestein and odd-length real paths;
 *   - counter-based input generation (splitmix64), identical to the oracle's.
 */
#include "fk_device.h"

using namespace fk;

namespace {

struct stw_desc {
    int nrad;
    int rad[12];
};

template <typename R>
__global__ void k_gen_stage_tw(cx<R> *out, stw_desc d, int total)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    /* locate stage s >= 1 owning entry t: stage s holds Ns*(Rs-1) entries
       laid out [k][r-1] */
    int off = 0, ns = d.rad[0];
    for (int s = 1; s < d.nrad; ++s) {
        int rs = d.rad[s], cnt = ns * (rs - 1);
        if (t < off + cnt) {
            int e = t - off, k = e / (rs - 1), r = e % (rs - 1) + 1;
            double c, sn;
            cexp_fp64((long long)r * k, (long long)ns * rs, c, sn);
            out[t] = mk((R)c, (R)sn);
            return;
        }
        off += cnt;
        ns *= rs;
    }
}

__global__ void k_gen_tw2(double *lo, double *hi, long long N, long long b)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long nhi = N / b + 1;
    double c, s;
    if (t < b) {
        cexp_fp64(t, N, c, s);
        lo[2 * t] = c;
        lo[2 * t + 1] = s;
    }
    if (t < nhi) {
        cexp_fp64(t * b, N, c, s);
        hi[2 * t] = c;
        hi[2 * t + 1] = s;
    }
}

template <typename R>
__global__ void k_gen_line_tw(cx<R> *out, long long n, long long count)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    double c, s;
    cexp_fp64(t, n, c, s);
    out[t] = mk((R)c, (R)s);
}

/* c_m = exp(-i pi m^2 / n) = W_2n^(m^2 mod 2n) */
template <typename R>
__global__ void k_gen_chirp(cx<R> *out, long long n)
{
    long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    long long e = (long long)(((unsigned __int128)m * (unsigned long long)m) % (unsigned long long)(2 * n));
    double c, s;
    cexp_fp64(e, 2 * n, c, s);
    out[m] = mk((R)c, (R)s);
}

/* b_m = conj(c_m) for |m| < n on a length-M circle, zero elsewhere (FP64) */
__global__ void k_gen_blue_b(cx<double> *out, long long n, long long M)
{
    long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= M) return;
    long long m = -1;
    if (j < n) m = j;
    else if (M - j < n) m = M - j;
    if (m < 0) { out[j] = mk(0.0, 0.0); return; }
    long long e = (long long)(((unsigned __int128)m * (unsigned long long)m) % (unsigned long long)(2 * n));
    double c, s;
    cexp_fp64(e, 2 * n, c, s);
    out[j] = mk(c, -s);
}

template <typename R>
__global__ void k_scale_round(cx<R> *out, const cx<double> *in, long long count, double scale)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    out[t] = mk((R)(in[t].x * scale), (R)(in[t].y * scale));
}

__device__ __forceinline__ unsigned long long smix(unsigned long long z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double u01(unsigned long long seed, unsigned long long stream,
                                      unsigned long long i)
{
    return (double)(smix(seed ^ (stream << 40) ^ i) >> 11) * (1.0 / 9007199254740992.0);
}

template <typename R>
__global__ void k_fill_uniform(R *out, long long count, unsigned long long seed,
                               unsigned long long stream)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long step = (long long)gridDim.x * blockDim.x;
    for (; t < count; t += step) out[t] = (R)(2.0 * u01(seed, stream, t) - 1.0);
}

template <typename R>
__global__ void k_fill_gauss(R *out, long long count, unsigned long long seed,
                             unsigned long long stream)
{
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long step = (long long)gridDim.x * blockDim.x;
    for (long long i = 2 * t; i < count; i += 2 * step) {
        double a = u01(seed, stream, i), b = u01(seed, stream, i + 1);
        double r = sqrt(-2.0 * log(1.0 - a));
        double sn, cs;
        sincospi(2.0 * b, &sn, &cs);
        out[i] = (R)(r * cs);
        if (i + 1 < count) out[i + 1] = (R)(r * sn);
    }
}

/* Thread mapping of the elementwise passes.  Lines of n > 128 elements:
   each workgroup covers a segment of EW_BS elements of one line (spl
   segments per line), so the line and its batch offsets are the same for
   the whole workgroup.  Shorter lines: lpb whole lines per workgroup.
   Blocks start at line line0. */
constexpr int EW_BS = 256;

template <typename R>
__global__ void __launch_bounds__(EW_BS) k_ew(fk_ew_args a, int lpb, long long spl, long long line0)
{
    long long line, jj;
    if (lpb > 1) {
        const unsigned nn = (unsigned)a.n, li = threadIdx.x / nn;
        if (li >= (unsigned)lpb) return;
        line = line0 + (long long)blockIdx.x * lpb + li;
        jj = (long long)(threadIdx.x - li * nn);
    } else {
        const long long b = blockIdx.x, ls = b / spl;
        line = line0 + ls;
        jj = (b - ls * spl) * EW_BS + threadIdx.x;
        if (jj >= a.n) return;
    }
    if (line >= a.nlines) return;
    const long long j = jj;
    long long i0, i1, i2;
    if (a.nlines <= 0xffffffffLL) {
        const unsigned l = (unsigned)line, n0 = (unsigned)a.nb[0], n1 = (unsigned)a.nb[1];
        const unsigned q = l / n0;
        i0 = l - q * n0;
        i1 = q % n1;
        i2 = q / n1;
    } else {
        const long long q = line / a.nb[0];
        i0 = line - q * a.nb[0];
        i1 = q % a.nb[1];
        i2 = q / a.nb[1];
    }
    const long long bi = i0 * a.ids[0] + i1 * a.ids[1] + i2 * a.ids[2];
    const long long bo = i0 * a.ods[0] + i1 * a.ods[1] + i2 * a.ods[2];
    const bool swi = a.flags & FK_SWAP_IN, swo = a.flags & FK_SWAP_OUT;
    const cx<R> *ci = reinterpret_cast<const cx<R> *>(a.in) + bi;
    cx<R> *co = reinterpret_cast<cx<R> *>(a.out) + bo;
    const R *ri = reinterpret_cast<const R *>(a.in) + bi;
    R *ro = reinterpret_cast<R *>(a.out) + bo;
    const cx<R> *aux = reinterpret_cast<const cx<R> *>(a.aux);
    const cx<R> *aux2 = reinterpret_cast<const cx<R> *>(a.aux2);
    auto fin = [&](cx<R> v) {
        if (a.flags & FK_SCALE) { v.x *= (R)a.scale; v.y *= (R)a.scale; }
        return swo ? swp(v) : v;
    };
    switch (a.op) {
    case FK_EW_COPY: {
        cx<R> v = ci[j * a.is];
        if (swi) v = swp(v);
        co[j * a.os] = fin(v);
        break;
    }
    case FK_EW_COPY_R: {
        R v = ri[j * a.is];
        if (a.flags & FK_SCALE) v *= (R)a.scale;
        ro[j * a.os] = v;
        break;
    }
    case FK_EW_SPLIT_GATHER: {
        const R *ri2 = reinterpret_cast<const R *>(a.in2) + bi;
        co[j * a.os] = mk(ri[j * a.is], ri2[j * a.is]);
        break;
    }
    case FK_EW_SPLIT_SCATTER: {
        R *ro2 = reinterpret_cast<R *>(a.out2) + bo;
        const cx<R> v = ci[j * a.is];
        ro[j * a.os] = v.x;
        ro2[j * a.os] = v.y;
        break;
    }
    case FK_EW_BS_PREP: {
        cx<R> v = mk((R)0, (R)0);
        if (j < a.p0) {
            v = ci[j * a.is];
            if (swi) v = swp(v);
            v = mul(v, aux[j]);
        }
        co[j * a.os] = v;
        break;
    }
    case FK_EW_BS_MUL:
        co[j * a.os] = swp(mul(ci[j * a.is], aux2[j]));
        break;
    case FK_EW_BS_POST:
        if (j < a.p0) co[j * a.os] = fin(mul(swp(ci[j * a.is]), aux[j]));
        break;
    case FK_EW_R2C_POST: {
        const long long L = a.p0;
        const cx<R> zk = ci[(j == L ? 0 : j) * a.is];
        const cx<R> zc = ci[(j == 0 ? 0 : L - j) * a.is];
        const R er = (R)0.5 * (zk.x + zc.x), ei = (R)0.5 * (zk.y - zc.y);
        const R orr = (R)0.5 * (zk.y + zc.y), oi = (R)-0.5 * (zk.x - zc.x);
        const cx<R> w = aux[j];
        co[j * a.os] = fin(mk(er + orr * w.x - oi * w.y, ei + orr * w.y + oi * w.x));
        break;
    }
    case FK_EW_C2R_PRE: {
        const long long L = a.p0;
        cx<R> xk = ci[j * a.is], xc = ci[(L - j) * a.is];
        if (j == 0) { xk.y = 0; xc.y = 0; }
        const cx<R> e = mk(xk.x + xc.x, xk.y - xc.y);
        const cx<R> d = mk(xk.x - xc.x, xk.y + xc.y);
        const cx<R> w = aux[j];
        const cx<R> o = mk(d.x * w.x + d.y * w.y, d.y * w.x - d.x * w.y);
        co[j * a.os] = mk(e.x - o.y, e.y + o.x);
        break;
    }
    case FK_EW_C2R_OUT: {
        cx<R> v = ci[j * a.is];
        if (a.flags & FK_SCALE) { v.x *= (R)a.scale; v.y *= (R)a.scale; }
        ro[(2 * j) * a.os] = v.x;
        ro[(2 * j + 1) * a.os] = v.y;
        break;
    }
    case FK_EW_WIDEN:
        co[j * a.os] = mk(ri[j * a.is], (R)0);
        break;
    case FK_EW_HERM: {
        const long long n = a.p0, h = n / 2;
        cx<R> v;
        if (j <= h) {
            v = ci[j * a.is];
            if (j == 0 || 2 * j == n) v.y = 0;
        } else {
            v = ci[(n - j) * a.is];
            v.y = -v.y;
        }
        co[j * a.os] = v;
        break;
    }
    case FK_EW_REAL: {
        R v = ci[j * a.is].x;
        if (a.flags & FK_SCALE) v *= (R)a.scale;
        ro[j * a.os] = v;
        break;
    }
    /* r2r pre / post passes; n = p0, sine kinds p1 = 1, aux[m] = exp(-i pi m / 4n) */
    case FK_EW_HC_POST: {
        /* bins C_k, k <= n/2 -> halfcomplex r_0..r_{n/2}, i_{(n+1)/2-1}..i_1
           (doc/reference.texi:917-925), or DHT Re C_k - Im C_k (C_{n-k} =
           conj C_k gives Re + Im above n/2) */
        const long long n = a.p0;
        const bool lo = 2 * j <= n;
        const cx<R> c = ci[(lo ? j : n - j) * a.is];
        R v;
        if (a.p1) v = lo ? c.x - c.y : c.x + c.y;
        else v = lo ? c.x : c.y;
        ro[j * a.os] = v;
        break;
    }
    case FK_EW_HC_PRE: {
        /* halfcomplex -> Hermitian bins k <= n/2 (no imaginary part at DC
           or, for even n, at n/2) */
        const long long n = a.p0;
        const R re = ri[j * a.is];
        const R im = (j == 0 || 2 * j == n) ? (R)0 : ri[(n - j) * a.is];
        co[j * a.os] = mk(re, im);
        break;
    }
    case FK_EW_DCT2_PRE: {
        /* v_j = x_2j (j < ceil(n/2)), v_{n-1-j} = x_{2j+1}; the sine kind
           negates odd samples (DST-II_k(x) = DCT-II_{n-1-k}((-1)^j x_j)) */
        const long long n = a.p0, h = (n + 1) / 2;
        const long long q = j < h ? 2 * j : 2 * (n - 1 - j) + 1;
        R v = ri[q * a.is];
        if (a.p1 && (q & 1)) v = -v;
        ro[j * a.os] = v;
        break;
    }
    case FK_EW_DCT2_POST: {
        /* D_k = 2 Re(exp(-i pi k / 2n) V_k), V_{n-k} = conj V_k */
        const long long n = a.p0, k = j;
        const bool lo = 2 * k <= n;
        cx<R> V = ci[(lo ? k : n - k) * a.is];
        if (!lo) V.y = -V.y;
        const cx<R> w = aux[2 * k];
        ro[(a.p1 ? n - 1 - k : k) * a.os] = (R)2 * (w.x * V.x - w.y * V.y);
        break;
    }
    case FK_EW_DCT3_PRE: {
        /* V_k = exp(+i pi k / 2n) (x_k - i x_{n-k}), x_n = 0, k <= n/2; the
           sine kind reads x reversed (DST-III_k(x) = (-1)^k DCT-III_k(x')) */
        const long long n = a.p0, k = j;
        const long long q0 = a.p1 ? n - 1 - k : k, q1 = a.p1 ? k - 1 : n - k;
        const R xk = ri[q0 * a.is];
        const R xn = k == 0 ? (R)0 : ri[q1 * a.is];
        const cx<R> w = aux[2 * k];
        co[j * a.os] = mk(w.x * xk - w.y * xn, -(w.x * xn + w.y * xk));
        break;
    }
    case FK_EW_DCT3_POST: {
        /* y_2i = v_i, y_2i+1 = v_{n-1-i} */
        const long long n = a.p0;
        const long long q = (j & 1) ? n - 1 - (j >> 1) : (j >> 1);
        R v = ri[q * a.is];
        if (a.p1 && (j & 1)) v = -v;
        ro[j * a.os] = v;
        break;
    }
    case FK_EW_DCT4_PRE: {
        /* even n: z_m = (x_2m + i x_{n-1-2m}) exp(-i pi (4m+1) / 4n), m < n/2;
           the sine kind reads x reversed (DST-IV_k(x) = (-1)^k DCT-IV_k(x')) */
        const long long n = a.p0, m = j;
        const long long q0 = 2 * m, q1 = n - 1 - 2 * m;
        const R xa = ri[(a.p1 ? n - 1 - q0 : q0) * a.is];
        const R xb = ri[(a.p1 ? n - 1 - q1 : q1) * a.is];
        co[j * a.os] = mul(mk(xa, xb), aux[4 * m + 1]);
        break;
    }
    case FK_EW_DCT4_POST: {
        /* c_p = Z_p exp(-i pi p / n): Y_2p = 2 Re c_p, Y_{n-1-2p} = -2 Im c_p */
        const long long n = a.p0;
        const bool ev = !(j & 1);
        const long long p = ev ? (j >> 1) : ((n - 1 - j) >> 1);
        const cx<R> c = mul(ci[p * a.is], aux[4 * p]);
        R y = ev ? (R)2 * c.x : (R)-2 * c.y;
        if (a.p1 && !ev) y = -y;
        ro[j * a.os] = y;
        break;
    }
    case FK_EW_DCT4O_PRE: {
        /* odd n: a_j = x_j exp(-i pi j / 2n) for j < n, zero up to 2n */
        const long long n = a.p0;
        cx<R> v = mk((R)0, (R)0);
        if (j < n) {
            const R x = ri[(a.p1 ? n - 1 - j : j) * a.is];
            const cx<R> w = aux[2 * j];
            v = mk(x * w.x, x * w.y);
        }
        co[j * a.os] = v;
        break;
    }
    case FK_EW_DCT4O_POST: {
        /* Y_k = 2 Re(exp(-i pi (2k+1) / 4n) A_k), A = 2n-point DFT of a */
        const cx<R> c = mul(ci[j * a.is], aux[2 * j + 1]);
        R y = (R)2 * c.x;
        if (a.p1 && (j & 1)) y = -y;
        ro[j * a.os] = y;
        break;
    }
    case FK_EW_SYM_PRE: {
        /* even extension x_0..x_{n-1}, x_{n-2}..x_1 (length 2(n-1)), or odd
           extension 0, x_0..x_{n-1}, 0, -x_{n-1}..-x_0 (length 2(n+1)) */
        const long long n = a.p0;
        R v;
        if (!a.p1) {
            v = ri[(j < n ? j : 2 * (n - 1) - j) * a.is];
        } else if (j == 0 || j == n + 1) {
            v = (R)0;
        } else if (j <= n) {
            v = ri[(j - 1) * a.is];
        } else {
            v = -ri[(2 * (n + 1) - j - 1) * a.is];
        }
        ro[j * a.os] = v;
        break;
    }
    case FK_EW_SYM_POST:
        /* DCT-I: Y_k = Re B_k; DST-I: Y_k = -Im B_{k+1} */
        ro[j * a.os] = a.p1 ? -ci[(j + 1) * a.is].y : ci[j * a.is].x;
        break;
    default:
        break;
    }
}

/* Streaming copy, the practical HBM ceiling the transform passes are
   measured against: 16-byte loads and stores, U of them in flight per
   thread, each workgroup on a contiguous chunk of U*256*16 bytes per
   iteration; NT selects non-temporal accesses (no cache allocation).
   n16 counts 16-byte units; a byte tail is copied by the last thread. */
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int U, bool NT>
__global__ void __launch_bounds__(256) k_stream_copy(const u32x4 *__restrict__ src,
                                                     u32x4 *__restrict__ dst, long long n16)
{
    const long long step = (long long)gridDim.x * 256 * U;
    for (long long base = (long long)blockIdx.x * 256 * U + threadIdx.x; base < n16; base += step) {
        u32x4 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = base + (long long)u * 256;
            if (i < n16) {
                if constexpr (NT) v[u] = __builtin_nontemporal_load(src + i);
                else v[u] = src[i];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = base + (long long)u * 256;
            if (i < n16) {
                if constexpr (NT) __builtin_nontemporal_store(v[u], dst + i);
                else dst[i] = v[u];
            }
        }
    }
}

__global__ void k_copy_tail(const unsigned char *src, unsigned char *dst, long long n)
{
    for (long long i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

inline unsigned nblocks(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

} // namespace

extern "C" {

int fkm_gen_stage_tw(void *out, int prec, int nrad, const int *rad, int total, hipStream_t st)
{
    if (total <= 0) return 0;
    stw_desc d;
    d.nrad = nrad;
    for (int i = 0; i < 12; ++i) d.rad[i] = i < nrad ? rad[i] : 1;
    if (prec == 0)
        hipLaunchKernelGGL(k_gen_stage_tw<double>, dim3(nblocks(total, 256)), dim3(256), 0, st,
                           (cx<double> *)out, d, total);
    else
        hipLaunchKernelGGL(k_gen_stage_tw<float>, dim3(nblocks(total, 256)), dim3(256), 0, st,
                           (cx<float> *)out, d, total);
    return (int)hipGetLastError();
}

int fkm_gen_tw2(double *lo, double *hi, long long N, long long b, hipStream_t st)
{
    long long cnt = b > N / b + 1 ? b : N / b + 1;
    hipLaunchKernelGGL(k_gen_tw2, dim3(nblocks(cnt, 256)), dim3(256), 0, st, lo, hi, N, b);
    return (int)hipGetLastError();
}

int fkm_gen_line_tw(void *out, int prec, long long n, long long count, hipStream_t st)
{
    if (count <= 0) return 0;
    if (prec == 0)
        hipLaunchKernelGGL(k_gen_line_tw<double>, dim3(nblocks(count, 256)), dim3(256), 0, st,
                           (cx<double> *)out, n, count);
    else
        hipLaunchKernelGGL(k_gen_line_tw<float>, dim3(nblocks(count, 256)), dim3(256), 0, st,
                           (cx<float> *)out, n, count);
    return (int)hipGetLastError();
}

int fkm_gen_chirp(void *out, int prec, long long n, hipStream_t st)
{
    if (prec == 0)
        hipLaunchKernelGGL(k_gen_chirp<double>, dim3(nblocks(n, 256)), dim3(256), 0, st,
                           (cx<double> *)out, n);
    else
        hipLaunchKernelGGL(k_gen_chirp<float>, dim3(nblocks(n, 256)), dim3(256), 0, st,
                           (cx<float> *)out, n);
    return (int)hipGetLastError();
}

int fkm_gen_blue_b(void *out, long long n, long long M, hipStream_t st)
{
    hipLaunchKernelGGL(k_gen_blue_b, dim3(nblocks(M, 256)), dim3(256), 0, st,
                       (cx<double> *)out, n, M);
    return (int)hipGetLastError();
}

int fkm_scale_round(void *out, int prec, const void *in, long long count, double scale,
                    hipStream_t st)
{
    if (prec == 0)
        hipLaunchKernelGGL(k_scale_round<double>, dim3(nblocks(count, 256)), dim3(256), 0, st,
                           (cx<double> *)out, (const cx<double> *)in, count, scale);
    else
        hipLaunchKernelGGL(k_scale_round<float>, dim3(nblocks(count, 256)), dim3(256), 0, st,
                           (cx<float> *)out, (const cx<double> *)in, count, scale);
    return (int)hipGetLastError();
}

int fkm_fill(void *out, int prec, long long count, unsigned long long seed,
             unsigned long long stream, int gauss, hipStream_t st)
{
    if (count <= 0) return 0;
    unsigned nb = nblocks(count, 256);
    if (nb > 65536) nb = 65536;
    if (prec == 0) {
        if (gauss)
            hipLaunchKernelGGL(k_fill_gauss<double>, dim3(nb), dim3(256), 0, st, (double *)out,
                               count, seed, stream);
        else
            hipLaunchKernelGGL(k_fill_uniform<double>, dim3(nb), dim3(256), 0, st, (double *)out,
                               count, seed, stream);
    } else {
        if (gauss)
            hipLaunchKernelGGL(k_fill_gauss<float>, dim3(nb), dim3(256), 0, st, (float *)out,
                               count, seed, stream);
        else
            hipLaunchKernelGGL(k_fill_uniform<float>, dim3(nb), dim3(256), 0, st, (float *)out,
                               count, seed, stream);
    }
    return (int)hipGetLastError();
}

/* Streaming copy of `bytes` (16-byte aligned pointers); mode 0: plain
   16-byte accesses, 1: non-temporal.  The grid covers every CU several
   times over (ctas_per_cu workgroups per CU) and loops over the rest. */
int fkm_stream_copy(void *dst, const void *src, long long bytes, int mode, int ctas, hipStream_t st)
{
    const long long n16 = bytes / 16, tail = bytes - n16 * 16;
    if (((uintptr_t)dst | (uintptr_t)src) % 16) return (int)hipErrorInvalidValue;
    if (ctas <= 0) ctas = 2048;
    long long need = (n16 + 256 * 4 - 1) / (256 * 4);
    unsigned grid = (unsigned)(need < ctas ? (need > 0 ? need : 1) : ctas);
    if (n16 > 0) {
        if (mode == 1)
            hipLaunchKernelGGL((k_stream_copy<4, true>), dim3(grid), dim3(256), 0, st,
                               (const u32x4 *)src, (u32x4 *)dst, n16);
        else
            hipLaunchKernelGGL((k_stream_copy<4, false>), dim3(grid), dim3(256), 0, st,
                               (const u32x4 *)src, (u32x4 *)dst, n16);
    }
    if (tail)
        hipLaunchKernelGGL(k_copy_tail, dim3(1), dim3(64), 0, st,
                           (const unsigned char *)src + n16 * 16, (unsigned char *)dst + n16 * 16,
                           tail);
    return (int)hipGetLastError();
}

int fkm_ew(int prec, const fk_ew_args *a, hipStream_t st)
{
    if (a->nlines <= 0 || a->n <= 0) return 0;
    const int lpb = a->n <= EW_BS / 2 ? (int)(EW_BS / a->n) : 1;
    const long long spl = lpb > 1 ? 1 : (a->n + EW_BS - 1) / EW_BS;
    const long long total = lpb > 1 ? (a->nlines + lpb - 1) / lpb : a->nlines * spl;
    long long chunk = ((1LL << 30) / spl) * spl;        /* whole lines per launch */
    if (chunk < spl) chunk = spl;
    for (long long b0 = 0; b0 < total; b0 += chunk) {
        const long long nb = total - b0 < chunk ? total - b0 : chunk;
        const long long line0 = lpb > 1 ? b0 * lpb : b0 / spl;
        if (prec == 0)
            hipLaunchKernelGGL(k_ew<double>, dim3((unsigned)nb), dim3(EW_BS), 0, st, *a, lpb, spl,
                               line0);
        else
            hipLaunchKernelGGL(k_ew<float>, dim3((unsigned)nb), dim3(EW_BS), 0, st, *a, lpb, spl,
                               line0);
    }
    return (int)hipGetLastError();
}

} // extern "C"
