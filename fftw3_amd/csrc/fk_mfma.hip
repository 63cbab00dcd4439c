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

/*
 * NL lines of length L = 16^S per workgroup of NT threads, staged through
 * LDS: a coalesced load of the whole tile (threads along the line for unit
 * element stride, across lines otherwise), S radix-16 stages on the matrix
 * cores with the tile in LDS, and a coalesced store of the tile with the
 * output post-op (four-step twiddle, scale, backward swap).  Each wave runs
 * groups of 32 butterflies (one MFMA column each) of the tile's
 * NL * L/16 butterflies per stage.
 */
template <int S, int NL, int NT>
__global__ void __launch_bounds__(NT) k_mfma(fk_args a)
{
    constexpr int L = S == 1 ? 16 : S == 2 ? 256 : 4096;
    constexpr int NB = L / 16;                /* butterflies per line and stage */
    constexpr int W = NT / 64;                /* waves */
    constexpr int NG = NL * NB / 32;          /* groups of 32 butterflies */
    constexpr int GPW = (NG + W - 1) / W;     /* groups per wave */
    constexpr int LP = L + L / 16 + 2;        /* LDS pitch of a line */
    static_assert((NL * NB) % 32 == 0, "tile must hold whole MFMA groups");
    static_assert(NT % NL == 0 && NT % 64 == 0, "thread layout");
    constexpr long long NONE = -0x7fffffffffffffffLL - 1;
    __shared__ cx<float> lds[NL * LP];
    __shared__ long long lin[NL], lout[NL], li0[NL];
    auto idx = [](int lt, int j) { return lt * LP + j + (j >> 4); };

    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int n = lane & 31, h = lane >> 5;
    /* XCD-contiguous tile order: neighbouring tiles (which share cache
       lines when the pass is strided) run on one XCD and its L2 */
    const long long tile = fk_remap_line(a.xcd_remap);
    if (tid < NL) {
        const long long line = tile * NL + tid;
        const bool act = line < a.nlines;
        long long q = act ? line : 0;
        const long long i0 = q % a.nb[0];
        q /= a.nb[0];
        const long long i1 = q % a.nb[1], i2 = q / a.nb[1];
        /* offsets may be negative (negative strides): inactive lines are
           marked with the one value no offset takes */
        lin[tid] = act ? i0 * a.ids[0] + i1 * a.ids[1] + i2 * a.ids[2] : NONE;
        lout[tid] = act ? i0 * a.ods[0] + i1 * a.ods[1] + i2 * a.ods[2] : NONE;
        li0[tid] = i0;
    }
    __syncthreads();
    const cx<float> *in = reinterpret_cast<const cx<float> *>(a.in);
    cx<float> *out = reinterpret_cast<cx<float> *>(a.out);
    const cx<float> *tw = reinterpret_cast<const cx<float> *>(a.tw);
    const bool swi = a.flags & FK_SWAP_IN, swo = a.flags & FK_SWAP_OUT;

    /* load the tile */
    if (a.is == 1) {
        for (int e = tid; e < NL * L; e += NT) {
            const int lt = e / L, j = e % L;
            const long long base = lin[lt];
            cx<float> v = base != NONE ? in[base + j] : mk(0.f, 0.f);
            lds[idx(lt, j)] = swi ? swp(v) : v;
        }
    } else {
        for (int e = tid; e < NL * L; e += NT) {
            const int lt = e % NL, j = e / NL;
            const long long base = lin[lt];
            cx<float> v = base != NONE ? in[base + (long long)j * a.is] : mk(0.f, 0.f);
            lds[idx(lt, j)] = swi ? swp(v) : v;
        }
    }

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
    __syncthreads();

#pragma unroll
    for (int st = 0; st < S; ++st) {
        const int Ns = st == 0 ? 1 : st == 1 ? 16 : 256;
        const int toff = st == 1 ? 0 : 16 * 15;
        cx<float> x[GPW][8];
        /* inputs r = 4s + 2h + t of this lane's butterfly in each group */
#pragma unroll
        for (int gi = 0; gi < GPW; ++gi) {
            const int g = wave + gi * W;
            if (g < NG) {
                const int id = g * 32 + n, lt = id / NB, b = id % NB, k = b % Ns;
#pragma unroll
                for (int s = 0; s < 4; ++s)
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const int r = 4 * s + 2 * h + t;
                        cx<float> v = lds[idx(lt, b + r * NB)];
                        if (st > 0 && r) v = mul(v, tw[toff + k * 15 + (r - 1)]);
                        x[gi][2 * s + t] = v;
                    }
            }
        }
        /* every lane of the workgroup has read this stage's inputs */
        __syncthreads();
#pragma unroll
        for (int gi = 0; gi < GPW; ++gi) {
            const int g = wave + gi * W;
            if (g < NG) {
                const int id = g * 32 + n, lt = id / NB, b = id % NB, k = b % Ns;
                /* per-column power-of-two scale (the column lives in lanes
                   n and n+32) */
                float m = 0.f;
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    m = fmaxf(m, fmaxf(fabsf(x[gi][i].x), fabsf(x[gi][i].y)));
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
                    split(x[gi][2 * s].x * sc, hh, ll); Bhi[0] = hh; Blo[0] = ll;
                    split(x[gi][2 * s].y * sc, hh, ll); Bhi[1] = hh; Blo[1] = ll;
                    split(x[gi][2 * s + 1].x * sc, hh, ll); Bhi[2] = hh; Blo[2] = ll;
                    split(x[gi][2 * s + 1].y * sc, hh, ll); Bhi[3] = hh; Blo[3] = ll;
                    acc = __builtin_amdgcn_mfma_f32_32x32x8f16(Ahi[s], Bhi, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x8f16(Ahi[s], Blo, acc, 0, 0, 0);
                    acc = __builtin_amdgcn_mfma_f32_32x32x8f16(Alo[s], Bhi, acc, 0, 0, 0);
                }
                /* outputs: rows 8j + 4h + u, u = 0..3 -> complex kk = 4j + 2h + u/2 */
                const int base = (b - k) * 16 + k;
#pragma unroll
                for (int jj = 0; jj < 4; ++jj)
#pragma unroll
                    for (int t = 0; t < 2; ++t) {
                        const int kk = 4 * jj + 2 * h + t;
                        lds[idx(lt, base + kk * Ns)] =
                            mk(acc[4 * jj + 2 * t] * usc, acc[4 * jj + 2 * t + 1] * usc);
                    }
            }
        }
        __syncthreads();
    }

    /* store the tile with the output post-op */
    const bool twout = a.flags & FK_TW_OUT, scl = a.flags & FK_SCALE;
    if (a.os == 1) {
        for (int e = tid; e < NL * L; e += NT) {
            const int lt = e / L, k = e % L;
            const long long base = lout[lt];
            if (base == NONE) continue;
            cx<float> y = lds[idx(lt, k)];
            if (twout) y = mul(y, tw2<float>(a, (long long)k * li0[lt]));
            if (scl) { y.x *= (float)a.scale; y.y *= (float)a.scale; }
            out[base + k] = swo ? swp(y) : y;
        }
    } else {
        /* lines fastest: this thread keeps line lt and steps k by NT/NL, so
           the four-step twiddle W_N^(k*i0) is looked up in the FP64 table
           every 16th element and stepped by W_N^((NT/NL)*i0) in FP64 in
           between (at most 15 FP64 products from a table value) */
        constexpr int DK = NT / NL;
        const int lt = tid % NL;
        const long long base = lout[lt], i0 = li0[lt];
        cx<double> wcur = mk(1.0, 0.0), wstep = mk(1.0, 0.0);
        if (twout) wstep = tw2d(a, (long long)DK * i0);
        int it = 0;
        for (int k = tid / NL; k < L; k += DK, ++it) {
            if (base == NONE) break;
            cx<float> y = lds[idx(lt, k)];
            if (twout) {
                wcur = (it % 16 == 0) ? tw2d(a, (long long)k * i0) : mul(wcur, wstep);
                y = mul(y, mk((float)wcur.x, (float)wcur.y));
            }
            if (scl) { y.x *= (float)a.scale; y.y *= (float)a.scale; }
            out[base + (long long)k * a.os] = swo ? swp(y) : y;
        }
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
    /* tiles: 64 lines of 16 (128 threads), 16 lines of 256 (256
       threads), 2 lines of 4096 (512 threads); block = tpl * nl */
    k->nl = S == 1 ? 64 : S == 2 ? 16 : 2;
    k->tpl = (S == 1 ? 128 : S == 2 ? 256 : 512) / k->nl;
    k->mode = FK_ROW;
    k->io = FK_IO_PLAIN;
    k->lds = 0;
    k->twsz = S == 1 ? 0 : S == 2 ? 16 * 15 : 16 * 15 + 256 * 15;
    k->fn = S == 1 ? (const void *)&k_mfma<1, 64, 128>
          : S == 2 ? (const void *)&k_mfma<2, 16, 256>
                   : (const void *)&k_mfma<3, 2, 512>;
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
