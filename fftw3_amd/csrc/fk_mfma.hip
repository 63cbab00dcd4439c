/*
 * Split-precision matrix-core variant of the radix-16 Stockham pass
 * (FP32 data, L = 16^S, S = 1..3).
 *
 * Every radix-16 butterfly of a stage is one column of a real 32x32 by 32
 * matrix product: the 16-point DFT F (F[k][r] = exp(-2*pi*i*k*r/16)) written
 * as a real matrix over interleaved (re, im) vectors,
 *     y[2k]   = sum_r Re F[k][r] x[2r] - Im F[k][r] x[2r+1]
 *     y[2k+1] = sum_r Im F[k][r] x[2r] + Re F[k][r] x[2r+1].
 * A wave holds 32 butterflies (columns) and runs
 * v_mfma_f32_32x32x8_f16 over 4 K-steps.  Operands are split into FP16
 * hi + lo parts (x = hi + lo, lo = fp16(x - hi)) and accumulated in FP32 as
 * F_hi*x_hi + F_hi*x_lo + F_lo*x_hi.  Each column is scaled by an exact
 * power of two before the split (max |x| in [2^8, 2^9)) so FP16 neither
 * overflows nor loses lo bits to underflow; the output is scaled back.
 *
 * Operand layouts of v_mfma_f32_32x32x8_f16 (wave64, lane l):
 *   A (32 x 8): lane l holds A[l % 32][4*(l/32) + q], q = 0..3
 *   B (8 x 32): lane l holds B[4*(l/32) + q][l % 32]
 *   D (32x32):  lane l, register v holds D[8*(v/4) + 4*(l/32) + v%4][l % 32]
 * (checked on the device by tests/test_gpu_precision.py::test_mfma_layout).
 *
 * Line addressing, twiddle tables (stages 16,16,16), backward-by-swap and
 * the fused four-step output twiddle are those of the Stockham kernels, so
 * this kernel can stand in for any pass of length 16^S.
 */
#include "fk_device.h"

using namespace fk;

typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

namespace {

__device__ __forceinline__ long long fk_remap_line(unsigned xcd)
{
    long long w = blockIdx.x;
    if (!xcd) return w;
    /* workgroups are dispatched round-robin over 8 XCDs */
    long long nwg = gridDim.x, per = nwg / 8, rem = nwg % 8;
    long long x = w % 8, i = w / 8;
    return x * per + (x < rem ? x : rem) + i;
}

/* 32x32 real DFT-16 matrix (row = output real index, col = input real index) */
struct Dft16Real {
    float v[32][32];
    constexpr Dft16Real() : v()
    {
        for (int k = 0; k < 16; ++k)
            for (int r = 0; r < 16; ++r) {
                float c = (float)kcos((long long)k * r, 16), s = (float)-ksin((long long)k * r, 16);
                v[2 * k][2 * r] = c;
                v[2 * k][2 * r + 1] = -s;
                v[2 * k + 1][2 * r] = s;
                v[2 * k + 1][2 * r + 1] = c;
            }
    }
};

__constant__ Dft16Real c_dft16 = Dft16Real();

__device__ __forceinline__ void split(float x, _Float16 &hi, _Float16 &lo)
{
    hi = (_Float16)x;
    lo = (_Float16)(x - (float)hi);
}

template <int S>
__global__ void __launch_bounds__(64 * ((S == 3) ? 8 : 1)) k_mfma(fk_args a)
{
    constexpr int L = S == 1 ? 16 : S == 2 ? 256 : 4096;
    constexpr int NB = L / 16;          /* butterflies per stage */
    constexpr int W = S == 3 ? 8 : 1;   /* waves per workgroup */
    __shared__ cx<float> lds[L + L / 16 + 2];
    auto idx = [](int j) { return j + (j >> 4); };

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int n = lane & 31, h = lane >> 5;
    /* XCD-contiguous line order: neighbouring lines (which share cache
       lines when the pass is strided) run on one XCD and its L2 */
    const long long line = fk_remap_line(a.xcd_remap);
    const bool act = line < a.nlines;
    long long q = act ? line : 0;
    const long long i0 = q % a.nb[0];
    q /= a.nb[0];
    const long long i1 = q % a.nb[1], i2 = q / a.nb[1];
    const long long bin = i0 * a.ids[0] + i1 * a.ids[1] + i2 * a.ids[2];
    const long long bout = i0 * a.ods[0] + i1 * a.ods[1] + i2 * a.ods[2];
    const cx<float> *in = reinterpret_cast<const cx<float> *>(a.in) + bin;
    cx<float> *out = reinterpret_cast<cx<float> *>(a.out);
    const cx<float> *tw = reinterpret_cast<const cx<float> *>(a.tw);
    const bool swi = a.flags & FK_SWAP_IN, swo = a.flags & FK_SWAP_OUT;

    /* A operand (DFT matrix), hi/lo, for the 4 K-steps */
    h4 Ahi[4], Alo[4];
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) {
            _Float16 hi, lo;
            split(c_dft16.v[n][8 * s + 4 * h + qq], hi, lo);
            Ahi[s][qq] = hi;
            Alo[s][qq] = lo;
        }

    const int b = wave * 32 + n;        /* this lane's butterfly */
    const bool bval = b < NB;
#pragma unroll
    for (int st = 0; st < S; ++st) {
        const int Ns = st == 0 ? 1 : st == 1 ? 16 : 256;
        const int k = bval ? b % Ns : 0;
        cx<float> x[8];
        /* inputs r = 4s + 2h + t */
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int r = 4 * s + 2 * h + t;
                const int j = b + r * NB;
                cx<float> v = mk(0.f, 0.f);
                if (bval) {
                    if (st == 0) {
                        if (act) v = in[j * a.is];
                        if (swi) v = swp(v);
                    } else {
                        v = lds[idx(j)];
                        if (r) v = mul(v, tw[(st == 1 ? 0 : 16 * 15) + k * 15 + (r - 1)]);
                    }
                }
                x[2 * s + t] = v;
            }
        /* per-column power-of-two scale (the column lives in lanes n, n+32) */
        float m = 0.f;
#pragma unroll
        for (int i = 0; i < 8; ++i) m = fmaxf(m, fmaxf(fabsf(x[i].x), fabsf(x[i].y)));
        m = fmaxf(m, __shfl_xor(m, 32));
        const int e = m > 0.f ? ilogbf(m) : 0;
        const float sc = ldexpf(1.f, 8 - e), usc = ldexpf(1.f, e - 8);
        f16v acc;
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] = 0.f;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            h4 Bhi, Blo;
            _Float16 hh, ll;
            split(x[2 * s].x * sc, hh, ll); Bhi[0] = hh; Blo[0] = ll;
            split(x[2 * s].y * sc, hh, ll); Bhi[1] = hh; Blo[1] = ll;
            split(x[2 * s + 1].x * sc, hh, ll); Bhi[2] = hh; Blo[2] = ll;
            split(x[2 * s + 1].y * sc, hh, ll); Bhi[3] = hh; Blo[3] = ll;
            acc = __builtin_amdgcn_mfma_f32_32x32x8f16(Ahi[s], Bhi, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x8f16(Ahi[s], Blo, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x8f16(Alo[s], Bhi, acc, 0, 0, 0);
        }
        /* every lane of the workgroup has read this stage's inputs */
        if (st > 0) __syncthreads();
        /* outputs: rows 8j + 4h + u, u = 0..3 -> complex kk = 4j + 2h + u/2 */
        const int base = (b - k) * 16 + k;
#pragma unroll
        for (int jj = 0; jj < 4; ++jj)
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int kk = 4 * jj + 2 * h + t;
                cx<float> y = mk(acc[4 * jj + 2 * t] * usc, acc[4 * jj + 2 * t + 1] * usc);
                const int pos = base + kk * Ns;
                if (!bval) continue;
                if (st < S - 1) {
                    lds[idx(pos)] = y;
                } else if (act) {
                    if (a.flags & FK_TW_OUT) y = mul(y, tw2<float>(a, (long long)pos * i0));
                    if (a.flags & FK_SCALE) { y.x *= (float)a.scale; y.y *= (float)a.scale; }
                    out[bout + pos * a.os] = swo ? swp(y) : y;
                }
            }
        if (st < S - 1) __syncthreads();
    }
}

/* Layout probe: D = A * B with A[i][k] = i + 32k/64... values chosen so
   every (row, col) product is distinct; writes D row-major. */
__global__ void k_mfma_probe(float *out)
{
    const int lane = threadIdx.x;
    h4 A, B;
    for (int qq = 0; qq < 4; ++qq) {
        const int kk = 4 * (lane >> 5) + qq;
        A[qq] = (_Float16)((lane & 31) == kk ? 1.f : 0.f);          /* A[i][k] = (i == k) */
        B[qq] = (_Float16)(float)(kk * 32 + (lane & 31)) * (_Float16)(1.f / 256.f);
    }
    f16v acc;
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    acc = __builtin_amdgcn_mfma_f32_32x32x8f16(A, B, acc, 0, 0, 0);
    for (int v = 0; v < 16; ++v) {
        const int row = 8 * (v / 4) + 4 * (lane >> 5) + (v % 4);
        out[row * 32 + (lane & 31)] = acc[v];
    }
}

} // namespace

extern "C" {

/* descriptor of the matrix-core kernel for length 16^S, shaped like a
   catalog entry so the host can launch it as any pass */
int fkm_mfma_kernel(int L, fk_kernel *k)
{
    int S = L == 16 ? 1 : L == 256 ? 2 : L == 4096 ? 3 : 0;
    int i;
    if (!S) return -1;
    k->prec = 1;
    k->L = L;
    k->nrad = S;
    for (i = 0; i < 12; ++i) k->rad[i] = i < S ? 16 : 0;
    k->tpl = S == 3 ? 512 : 64;
    k->nl = 1;
    k->mode = FK_ROW;
    k->io = FK_IO_PLAIN;
    k->lds = 0;
    k->twsz = S == 1 ? 0 : S == 2 ? 16 * 15 : 16 * 15 + 256 * 15;
    k->fn = S == 1 ? (const void *)&k_mfma<1> : S == 2 ? (const void *)&k_mfma<2>
                                                     : (const void *)&k_mfma<3>;
    k->variant = 100;
    return 0;
}

/* runs the layout probe: out[32*32] = D of one 32x32x8 product whose
   expected value at (row, col) is B[row][col] for row < 8, 0 otherwise */
int fkm_mfma_probe(float *dev_out, hipStream_t st)
{
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, st, dev_out);
    return (int)hipGetLastError();
}

} // extern "C"
