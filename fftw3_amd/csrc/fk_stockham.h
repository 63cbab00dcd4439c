/*
 * LDS-resident Stockham kernel: every workgroup transforms NL lines of
 * length L (compile-time L, radix sequence, thread layout).
 *
 * Stage s with radix Rs and Ns = R0*...*R(s-1) (autosort Stockham,
 * decimation in time):
 *   butterfly b in [0, L/Rs), k = b mod Ns
 *   inputs   x[b + r*L/Rs]                        r = 0..Rs-1
 *   twiddle  x_r *= W_{Ns*Rs}^{r*k}
 *   outputs  y[(b/Ns)*Ns*Rs + k + r*Ns] = DFT_Rs(x)_r
 * The first stage reads straight from HBM and the last stage writes straight
 * to HBM (both patterns are unit-stride across threads); the stages between
 * exchange through LDS.  This replaces FFTW's Cooley-Tukey recursion with
 * twiddle codelets (dft/ct.c:34-45, dft/dftw-direct.c:46-56) for every
 * length that fits one workgroup.
 *
 * Line addressing follows FFTW's tensor semantics (kernel/ifftw.h:417-432):
 * element j of line (i0,i1,i2) is at i0*ids0 + i1*ids1 + i2*ids2 + j*is.
 */
#ifndef FK_STOCKHAM_H
#define FK_STOCKHAM_H

#include "fk_device.h"

namespace fk {

template <int... RS>
struct RadList {
    static constexpr int n = sizeof...(RS);
    static constexpr int r[n] = {RS...};
    static constexpr int L()
    {
        int p = 1;
        for (int i = 0; i < n; ++i) p *= r[i];
        return p;
    }
    static constexpr int ns(int s)
    {
        int p = 1;
        for (int i = 0; i < s; ++i) p *= r[i];
        return p;
    }
    /* offset of stage s in the twiddle table (stage 0 has none) */
    static constexpr int twoff(int s)
    {
        int o = 0;
        for (int t = 1; t < s; ++t) o += ns(t) * (r[t] - 1);
        return o;
    }
    static constexpr int twsize() { return twoff(n); }
};

/* Element index b + OFF of a line, b a multiple of ALIGN.  The padded row
   layout keeps the compile-time part OFF out of the per-element address
   arithmetic (it becomes the LDS instruction's offset); every other
   accessor takes the plain index through the conversion. */
template <int OFF, int ALIGN = 1>
struct At {
    int b;
    __device__ __forceinline__ constexpr operator int() const { return b + OFF; }
};

/* XL = 1 adds one scratch line for threads beyond TPL*NL (a persistent
   workgroup running a geometry smaller than itself). */
template <typename R, int L, int TPL, int NL, int MODE, int PAD, int XL = 0>
struct Geo {
    /* LDS pitch of one line (ROW) — one pad slot every PAD elements, one
       extra slot for the Nyquist bin of the real transforms */
    static constexpr int LP = L + L / PAD + 2;
    static constexpr int NLP = NL + XL;
    __device__ __forceinline__ static int idx(int lt, int j)
    {
        if constexpr (MODE == FK_ROW)
            return lt * LP + j + (int)((unsigned)j / (unsigned)PAD);
        else
            return j * NLP + lt;
    }
    /* (b + OFF) / PAD = b / PAD + OFF / PAD when PAD divides OFF or b */
    template <int OFF, int ALIGN>
    __device__ __forceinline__ static int idx(int lt, At<OFF, ALIGN> j)
    {
        if constexpr (MODE == FK_ROW && (OFF % PAD == 0 || ALIGN % PAD == 0))
            return lt * LP + j.b + (int)((unsigned)j.b / (unsigned)PAD) + (OFF + OFF / PAD);
        else
            return idx(lt, (int)j);
    }
    static constexpr unsigned lds_elems()
    {
        return MODE == FK_ROW ? (unsigned)(NLP * LP) : (unsigned)((L + 2) * NLP);
    }
};

/* Workgroup barrier of the Stockham stages (a policy parameter so that a
   kernel may substitute its own). */
struct SyncAll {
    __device__ __forceinline__ static void sync() { __syncthreads(); }
};

/* Wave-level ordering for stages whose LDS exchange stays inside one wave
   (row mapping with every line's threads in one wave, lines in disjoint
   LDS rows): a wave's LDS accesses complete in issue order, so only the
   compiler must be kept from moving accesses across the exchange point.
   Waves of the workgroup then run their stages independently. */
struct SyncWave {
    __device__ __forceinline__ static void sync()
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
};

/*
 * One Stockham stage.  LoadF(j) -> cx<R> and StoreF(k, cx<R>) abstract the
 * data source and sink; IN_LDS/OUT_LDS say whether they touch LDS, which
 * decides the barriers.
 */
template <typename R, class RL, int S, int L, int TPL, bool IN_LDS, bool OUT_LDS,
          class SY = SyncAll, class LoadF, class StoreF>
__device__ __forceinline__ void stage(int u, const cx<R> *__restrict__ tw,
                                      LoadF &&ld, StoreF &&st)
{
    constexpr int Rs = RL::r[S];
    constexpr int Ns = RL::ns(S);
    constexpr int NB = L / Rs;
    constexpr int BPT = (NB + TPL - 1) / TPL;
    cx<R> v[BPT][Rs];
    sfor<0, BPT>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const int b = u + i * TPL;
        if ((NB % TPL == 0) || b < NB) {
            sfor<0, Rs>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                v[i][r] = ld(At<r * NB>{b});
            });
        }
    });
    if constexpr (IN_LDS && OUT_LDS) SY::sync();
    sfor<0, BPT>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const int b = u + i * TPL;
        if ((NB % TPL == 0) || b < NB) {
            /* the last stage has Ns = NB, so b < Ns there: its outputs are
               b + r*Ns, a compile-time offset from the thread's index */
            const int k = S == RL::n - 1 ? b : b % Ns;
            if constexpr (S > 0) {
                const cx<R> *w = tw + RL::twoff(S) + k * (Rs - 1);
                sfor<1, Rs>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    v[i][r] = mul(v[i][r], w[r - 1]);
                });
            }
            Dft<R, Rs, 1>::run(v[i]);
            const int base = (b - k) * Rs + k;
            /* stage 0 (Ns = 1) writes Rs consecutive outputs from b*Rs */
            sfor<0, Rs>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                st(At<r * Ns, (S == 0 ? Rs : 1)>{base}, v[i][r]);
            });
        }
    });
    if constexpr (OUT_LDS) SY::sync();
}

/* Runs stages [S, RL::n) ; stage S loads with `first`, the last stage stores
   with `last`, everything between goes through LDS. */
template <typename R, class RL, int S, int L, int TPL, bool FIRST_LDS,
          bool LAST_LDS, class SY = SyncAll, class LdsF, class FirstF, class LastF>
__device__ __forceinline__ void run_stages(int u, const cx<R> *tw, LdsF &&lds,
                                           FirstF &&first, LastF &&last)
{
    constexpr int n = RL::n;
    auto lds_ld = [&](auto j) { return lds(j); };
    auto lds_st = [&](auto k, cx<R> v) { lds(k) = v; };
    if constexpr (S == 0 && n == 1) {
        stage<R, RL, 0, L, TPL, FIRST_LDS, LAST_LDS, SY>(u, tw, first, last);
    } else if constexpr (S == 0) {
        stage<R, RL, 0, L, TPL, FIRST_LDS, true, SY>(u, tw, first, lds_st);
        run_stages<R, RL, 1, L, TPL, FIRST_LDS, LAST_LDS, SY>(u, tw, lds, first, last);
    } else if constexpr (S == n - 1) {
        stage<R, RL, S, L, TPL, true, LAST_LDS, SY>(u, tw, lds_ld, last);
    } else {
        stage<R, RL, S, L, TPL, true, true, SY>(u, tw, lds_ld, lds_st);
        run_stages<R, RL, S + 1, L, TPL, FIRST_LDS, LAST_LDS, SY>(u, tw, lds, first, last);
    }
}

__device__ __forceinline__ long long remap_tile(unsigned xcd)
{
    long long w = blockIdx.x;
    if (!xcd) return w;
    /* workgroups are dispatched round-robin over 8 XCDs; give each XCD a
       contiguous run of tiles so neighbouring tiles share its L2 */
    long long nwg = gridDim.x, per = nwg / 8, rem = nwg % 8;
    long long x = w % 8, i = w / 8;
    return x * per + (x < rem ? x : rem) + i;
}

/*
 * Body of one workgroup tile: lines [tile*NL, tile*NL + NL).  NT is the
 * workgroup size; when NT > TPL*NL (persistent workgroups shared by several
 * geometries) the extra threads work on a scratch LDS line and never touch
 * global memory, so every thread still meets every barrier.
 */
template <typename R, int NL, int MODE, int IO, int PAD, class RL, int TPL, int NT,
          class SY = SyncAll>
__device__ __forceinline__ void stockham_body(const fk_args &a, long long tile,
                                              unsigned char *smem_raw)
{
    constexpr int L = RL::L();
    constexpr int NR = TPL * NL;
    using G = Geo<R, L, TPL, NL, MODE, PAD>;
    static_assert(NT >= NR, "workgroup smaller than the geometry");
    cx<R> *lds = reinterpret_cast<cx<R> *>(smem_raw);
    long long *lbase = reinterpret_cast<long long *>(lds + G::lds_elems());
    /* threads beyond TPL*NL read and write a 64-slot trash area */
    cx<R> *trash = reinterpret_cast<cx<R> *>(lbase + ((NL + 1) & ~1));

    const int tid = threadIdx.x;
    const bool real = (NT == NR) || tid < NR;
    const int lt = !real ? 0 : (MODE == FK_ROW) ? tid / TPL : tid % NL;
    const int u = !real ? (tid - NR) % TPL : (MODE == FK_ROW) ? tid % TPL : tid / NL;
    const long long line = tile * NL + lt;
    const bool act = real && line < a.nlines;

    long long i0 = 0, i1 = 0, i2 = 0;
    {
        long long q = act ? line : 0;
        i0 = q % a.nb[0];
        q /= a.nb[0];
        i1 = q % a.nb[1];
        i2 = q / a.nb[1];
    }
    const long long bin = i0 * a.ids[0] + i1 * a.ids[1] + i2 * a.ids[2];
    const long long bout = i0 * a.ods[0] + i1 * a.ods[1] + i2 * a.ods[2];
    const cx<R> *tw = reinterpret_cast<const cx<R> *>(a.tw);
    const unsigned fl = a.flags;
    const bool swi = fl & FK_SWAP_IN, swo = fl & FK_SWAP_OUT;
    /* output scale: multiplying by 1 when unscaled is exact and costs less
       than selecting between scaled and unscaled values */
    const R sc = (fl & FK_SCALE) ? (R)a.scale : (R)1;

    auto lds_at = [&](auto j) -> cx<R> & {
        if constexpr (NT == NR) return lds[G::idx(lt, j)];
        else return real ? lds[G::idx(lt, j)] : trash[tid & 63];
    };

    /* output post-op: twiddle, scale, swap.  The four-step twiddle of
       output k = b + r*Ns (last stage, r = 0..R-1 stored in order) is
       W_N^(k*i0): looked up from the FP64 table every ANCH outputs and
       stepped by W_N^(Ns*i0) in FP64 in between, so a thread gathers 2
       table entries per ANCH outputs instead of 2 per output.  FP64 data
       re-anchors every 4 (at most 3 products from a table value); FP32
       data once per butterfly (at most 15 FP64 products, ~1e-15, far below
       FP32 rounding). */
    constexpr int NS_LAST = RL::ns(RL::n - 1);
    constexpr int ANCH = sizeof(R) == 8 ? 4 : 16;
    cx<double> tw_cur = mk(1.0, 0.0), tw_step = mk(1.0, 0.0);
    if (fl & FK_TW_OUT) tw_step = tw2d(a, (long long)NS_LAST * i0);
    auto post = [&](int k, cx<R> v) -> cx<R> {
        if (fl & FK_TW_OUT) {
            if ((k / NS_LAST) % ANCH == 0) tw_cur = tw2d(a, (long long)k * i0);
            else tw_cur = mul(tw_cur, tw_step);
            v = mul(v, mk((R)tw_cur.x, (R)tw_cur.y));
        }
        v.x *= sc; v.y *= sc;
        return swo ? swp(v) : v;
    };

    if constexpr (IO == FK_IO_PLAIN || IO == FK_IO_TOUT) {
        /* element j of the thread's line sits (j - u) strides from the
           thread's own element u; the stages pass j = u + (compile-time
           offset), so each access is one uniform product and one 64-bit
           add.  Inactive lanes read line 0 (in bounds) and store nothing. */
        const long long is = a.is, os = a.os;
        const cx<R> *inu = reinterpret_cast<const cx<R> *>(a.in) + bin + u * is;
        cx<R> *out = reinterpret_cast<cx<R> *>(a.out);
        auto gld = [&](int j) -> cx<R> {
            cx<R> v = inu[(long long)(j - u) * is];
            return swi ? swp(v) : v;
        };
        if constexpr (IO == FK_IO_PLAIN) {
            cx<R> *outu = out + bout + u * os;
            auto gst = [&](int k, cx<R> v) {
                if (act) outu[(long long)(k - u) * os] = post(k, v);
            };
            run_stages<R, RL, 0, L, TPL, false, false, SY>(u, tw, lds_at, gld, gst);
        } else {
            auto lst = [&](auto k, cx<R> v) { lds_at(k) = v; };
            /* output offsets may be negative (negative strides): inactive
               lines are marked with the one value no offset takes */
            constexpr long long NONE = -0x7fffffffffffffffLL - 1; /* INT64_MIN */
            if (u == 0 && real) lbase[lt] = act ? bout : NONE;
            run_stages<R, RL, 0, L, TPL, false, true, SY>(u, tw, lds_at, gld, lst);
            /* the store below reads every line of the tile */
            if constexpr (!std::is_same<SY, SyncAll>::value) __syncthreads();
            /* lines-fastest store: consecutive threads write consecutive
               lines of the tile (unit output batch stride) */
            for (int e = tid; e < NL * L; e += NT) {
                const int l2 = e % NL, k = e / NL;
                const long long ob = lbase[l2];
                if (ob != NONE) {
                    cx<R> v = lds[G::idx(l2, k)];
                    v.x *= sc; v.y *= sc;
                    out[ob + k * a.os] = swo ? swp(v) : v;
                }
            }
        }
    } else if constexpr (IO == FK_IO_R2C_FULL) {
        /* r2c of a length-L real line whose half length has no kernel (odd
           L): the full L-point DFT of (x_j, 0); bins k <= L/2 are stored */
        const R *inr = reinterpret_cast<const R *>(a.in) + bin;
        cx<R> *out = reinterpret_cast<cx<R> *>(a.out);
        auto gld = [&](int j) -> cx<R> {
            return act ? mk(inr[(long long)j * a.is], (R)0) : mk((R)0, (R)0);
        };
        auto gst = [&](int k, cx<R> v) {
            if (act && 2 * k <= L) {
                v.x *= sc; v.y *= sc;
                out[bout + (long long)k * a.os] = v;
            }
        };
        run_stages<R, RL, 0, L, TPL, false, false, SY>(u, tw, lds_at, gld, gst);
    } else if constexpr (IO == FK_IO_C2R_FULL) {
        /* c2r of length L from bins k <= L/2: X_{L-k} = conj X_k (Im X_0 and,
           for even L, Im X_{L/2} ignored); the inverse DFT as a forward DFT
           of swap(X), whose real output part is Im of the result */
        const cx<R> *in = reinterpret_cast<const cx<R> *>(a.in) + bin;
        R *outr = reinterpret_cast<R *>(a.out) + bout;
        auto gld = [&](int j) -> cx<R> {
            if (!act) return mk((R)0, (R)0);
            cx<R> v;
            if (2 * j <= L) {
                v = in[(long long)j * a.is];
                if (j == 0 || 2 * j == L) v.y = (R)0;
            } else {
                v = in[(long long)(L - j) * a.is];
                v.y = -v.y;
            }
            return swp(v);
        };
        auto gst = [&](int k, cx<R> v) {
            if (!act) return;
            R y = v.y;
            y *= sc;
            outr[(long long)k * a.os] = y;
        };
        run_stages<R, RL, 0, L, TPL, false, false, SY>(u, tw, lds_at, gld, gst);
    } else if constexpr (IO == FK_IO_DCT4) {
        /* REDFT11 of n = 2L reals: z_m = (x_2m + i x_{n-1-2m}) exp(-i pi (4m+1) / 4n),
           Z = L-point DFT, c_p = Z_p exp(-i pi p / n), y_2p = 2 Re c_p,
           y_{n-1-2p} = -2 Im c_p.  RODFT11_k(x) = (-1)^k REDFT11_k(x reversed). */
        const R *inr = reinterpret_cast<const R *>(a.in) + bin;
        R *outr = reinterpret_cast<R *>(a.out) + bout;
        const cx<R> *rt = reinterpret_cast<const cx<R> *>(a.rtw);
        const bool sn = a.rmode == FK_RM_DST4;
        constexpr long long N2 = 2LL * L;
        auto gld = [&](int m) -> cx<R> {
            if (!act) return mk((R)0, (R)0);
            const long long q0 = 2LL * m, q1 = N2 - 1 - 2LL * m;
            const R xa = inr[(sn ? N2 - 1 - q0 : q0) * a.is];
            const R xb = inr[(sn ? N2 - 1 - q1 : q1) * a.is];
            return mul(mk(xa, xb), rt[4 * m + 1]);
        };
        auto gst = [&](int p, cx<R> v) {
            if (!act) return;
            const cx<R> c = mul(v, rt[4 * p]);
            R ye = (R)2 * c.x, yo = (R)-2 * c.y;
            ye *= sc; yo *= sc;
            outr[(2LL * p) * a.os] = ye;
            outr[(N2 - 1 - 2LL * p) * a.os] = sn ? -yo : yo;
        };
        run_stages<R, RL, 0, L, TPL, false, false, SY>(u, tw, lds_at, gld, gst);
    } else if constexpr (IO == FK_IO_R2C || IO == FK_IO_R2C_RR || IO == FK_IO_R2C_XR) {
        /* real line of 2L samples viewed as L complex z_j = x_2j + i x_2j+1;
           after the L-point DFT Z, X[k] = E[k] + W_2L^k O[k] with
           E = (Z[k] + conj Z[L-k])/2, O = (Z[k] - conj Z[L-k])/(2i),
           k = 0..L (the role of rdft/ct-hc2c.c:197-210's HC2C_VIA_DFT). */
        const R *inr = reinterpret_cast<const R *>(a.in) + bin;
        cx<R> *out = reinterpret_cast<cx<R> *>(a.out);
        R *outr = reinterpret_cast<R *>(a.out) + bout;
        const cx<R> *wk = reinterpret_cast<const cx<R> *>(a.aux);
        const cx<R> *rt = reinterpret_cast<const cx<R> *>(a.rtw);
        /* fused real-to-real kind; a compile-time FK_RM_NONE in the plain
           r2c / c2r instantiations */
        const unsigned rm = (IO == FK_IO_R2C_RR || IO == FK_IO_C2R_RR) ? a.rmode : 0u;
        constexpr long long N2 = 2LL * L;
        /* REDFT10 / RODFT10 read the line reordered (Makhoul):
           v_m = x_2m for m < L, x_{2(n-1-m)+1} above; the sine kind
           negates the odd samples (RODFT10_k(x) = REDFT10_{n-1-k}((-1)^j x_j)) */
        auto vld = [&](int m) -> R {
            const long long q = m < L ? 2LL * m : 2 * (N2 - 1 - m) + 1;
            const R v = inr[q * a.is];
            return (rm == FK_RM_DST2 && m >= L) ? -v : v;
        };
        /* REDFT00 / RODFT00 read the even extension x_0..x_{n-1},
           x_{n-2}..x_1 (n = L+1) or the odd one 0, x_0..x_{n-1}, 0,
           -x_{n-1}..-x_0 (n = L-1) of the line */
        auto eld = [&](long long m) -> R {
            if (rm == FK_RM_E00) return inr[(m <= L ? m : N2 - m) * a.is];
            if (m == 0 || m == L) return (R)0;
            return m < L ? inr[(m - 1) * a.is] : -inr[(N2 - m - 1) * a.is];
        };
        auto gld = [&](int j) -> cx<R> {
            if (!act) return mk((R)0, (R)0);
            if constexpr (sizeof(R) == 4) {
                /* v_2j, v_2j+1 are x_4j, x_4j+2 (first half) or x_{2n-4j-1},
                   x_{2n-4j-3} (second half): one 16-byte quad each */
                if ((rm == FK_RM_DCT2 || rm == FK_RM_DST2) && (fl & FK_REAL_QUAD)) {
                    const float4 *q4 = reinterpret_cast<const float4 *>(inr);
                    if (2 * j + 1 < L) {
                        const float4 q = q4[j];
                        return mk(q.x, q.z);
                    }
                    if (2 * j >= L) {
                        const float4 q = q4[L - 1 - j];
                        return rm == FK_RM_DST2 ? mk(-q.w, -q.y) : mk(q.w, q.y);
                    }
                }
            }
            if (rm == FK_RM_DCT2 || rm == FK_RM_DST2) return mk(vld(2 * j), vld(2 * j + 1));
            if (rm == FK_RM_E00 || rm == FK_RM_O00) return mk(eld(2 * j), eld(2 * j + 1));
            if (fl & FK_REAL_UNIT) return reinterpret_cast<const cx<R> *>(inr)[j];
            return mk(inr[(2 * j) * a.is], inr[(2 * j + 1) * a.is]);
        };
        auto lst = [&](auto k, cx<R> v) { lds_at(k) = v; };
        /* bin k <= L/2 gives bins k and L-k from the same Z[k], Z[L-k] and
           one twiddle (W_2L^(L-k) = -conj W_2L^k); bins k = u + i*TPL in
           groups of 8 (compile-time count): the group's twiddle loads are
           in flight together */
        constexpr int HK = L / 2;
        constexpr int NK2 = (HK + 1 + TPL - 1) / TPL;
        constexpr int GK = 8;
        constexpr bool XR = IO == FK_IO_R2C_XR;
        /* few bins per thread: their twiddles are loaded before the stages,
           so the load latency hides behind the FFT instead of following it */
        constexpr bool PRE = NK2 <= 20;
        cx<R> wpre[PRE ? NK2 : 1];
        if constexpr (PRE) {
            sfor<0, NK2>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                const int k = u + i * TPL;
                wpre[i] = wk[k <= HK ? k : HK];
            });
        }
        /* cross-row form: twiddles of the first column stage, W_n0^(n1*p)
           (n1 = the tile's line-group index, uniform over the workgroup) */
        cx<R> wp[XR ? NL : 1];
        if constexpr (XR) {
            const long long n1 = tile % a.nb[1];
            sfor<0, NL>([&](auto pc) {
                constexpr int pp = decltype(pc)::value;
                wp[pp] = pp == 0 ? mk((R)1, (R)0) : tw2<R>(a, n1 * pp);
            });
        }
        run_stages<R, RL, 0, L, TPL, false, true, SY>(u, tw, lds_at, gld, lst);
        /* cross-row form: the bins stay in registers until every bin of the
           line has been formed, then replace Z in the line's LDS row */
        cx<R> xs[XR ? 2 * NK2 : 1];
        constexpr long long NONE = -0x7fffffffffffffffLL - 1; /* INT64_MIN */
        if constexpr (XR) {
            if (u == 0 && real) lbase[lt] = act ? bout : NONE;
        }
        /* bin k's outputs: the complex bin, or the real outputs of the
           fused real-to-real kind */
        auto emit = [&](int k, cx<R> X) {
            if (act && rm == FK_RM_NONE) {
                out[bout + (long long)k * a.os] = X;
            } else if (act && rm == FK_RM_E00) {
                outr[(long long)k * a.os] = X.x;             /* DCT-I: Re X_k */
            } else if (act && rm == FK_RM_O00) {
                if (k >= 1 && k < L) outr[(long long)(k - 1) * a.os] = -X.y; /* DST-I */
            } else if (act) {
                /* bin k gives two real outputs, k and n-k:
                   R2HC r_k / i_k; DHT Re -+ Im; REDFT10
                   c = exp(-i pi k / 2n) X: 2 Re c, -2 Im c */
                R lo, hi;
                if (rm == FK_RM_HC) {
                    lo = X.x;
                    hi = X.y;
                } else if (rm == FK_RM_DHT) {
                    lo = X.x - X.y;
                    hi = X.x + X.y;
                } else {
                    const cx<R> c = mul(X, rt[2 * k]);
                    lo = (R)2 * c.x;
                    hi = (R)-2 * c.y;
                }
                long long ka = k, kb = N2 - k;
                if (rm == FK_RM_DST2) {
                    ka = N2 - 1 - k;
                    kb = k - 1;
                }
                outr[ka * a.os] = lo;
                if (k != 0 && k != L) outr[kb * a.os] = hi;
            }
        };
        sfor<0, (NK2 + GK - 1) / GK>([&](auto gc) {
            constexpr int g = decltype(gc)::value;
            constexpr int cnt = (NK2 - g * GK) < GK ? (NK2 - g * GK) : GK;
            cx<R> w[cnt];
            sfor<0, cnt>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                const int k = u + (g * GK + i) * TPL;
                if constexpr (PRE) w[i] = wpre[g * GK + i];
                else if (k <= HK) w[i] = wk[k];
            });
            sfor<0, cnt>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                const int k = u + (g * GK + i) * TPL;
                if (k <= HK) {
                    const cx<R> zk = lds_at(k);
                    const cx<R> zc = lds_at(k == 0 ? 0 : L - k);
                    const R er = (R)0.5 * (zk.x + zc.x), ei = (R)0.5 * (zk.y - zc.y);
                    /* O = (Z[k] - conj Z[L-k]) / (2i) */
                    const R orr = (R)0.5 * (zk.y + zc.y), oi = (R)-0.5 * (zk.x - zc.x);
                    const R wx = w[i].x, wy = w[i].y;
                    /* X[k] = E + W O; X[L-k] = conj(E - W O), written as the
                       form of bin L-k on its own (E' = conj E, O' = conj O,
                       W' = -conj W) so that both round as that form does */
                    cx<R> X = mk(er + orr * wx - oi * wy, ei + orr * wy + oi * wx);
                    cx<R> Y = mk(er + orr * (-wx) - (-oi) * wy, (-ei) + orr * wy + (-oi) * (-wx));
                    X.x *= sc; X.y *= sc;
                    Y.x *= sc; Y.y *= sc;
                    if constexpr (XR) {
                        xs[2 * (g * GK + i)] = X;
                        xs[2 * (g * GK + i) + 1] = Y;
                    } else {
                        emit(k, X);
                        if (2 * k != L) emit(L - k, Y);
                    }
                }
            });
        });
        if constexpr (XR) {
            /* every thread of the line has read Z: the bins replace it */
            SY::sync();
            sfor<0, NK2>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                const int k = u + i * TPL;
                if (k <= HK) {
                    lds_at(k) = xs[2 * i];
                    if (2 * k != L) lds_at(L - k) = xs[2 * i + 1];
                }
            });
            __syncthreads();
            /* first stage of the column DFT over the tile's NL rows
               n = n1 + m*n2 (n2 = the line's batch index i0, n1 = i1; the
               planner sets nb[0] = NL, so a tile is one n1): for every bin k
               y_p = W_n0^(n1*p) sum_n2 X_n2[k] W_NL^(n2*p), n0 = a.tN,
               stored to the tile's output line p */
            for (int k = tid; k <= L; k += NT) {
                cx<R> v[NL];
                sfor<0, NL>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    v[q] = lds[G::idx(q, k)];
                });
                Dft<R, NL, 1>::run(v);
                sfor<0, NL>([&](auto pc) {
                    constexpr int pp = decltype(pc)::value;
                    const long long ob = lbase[pp];
                    if (ob != NONE) out[ob + (long long)k * a.os] = pp == 0 ? v[0] : mul(v[pp], wp[pp]);
                });
            }
        }
    } else if constexpr (IO == FK_IO_C2R || IO == FK_IO_C2R_RR || IO == FK_IO_C2R_XR) {
        /* half-length inverse: Z[k] = E[k] + i O[k],
           E[k] = X[k] + conj X[L-k], O[k] = (X[k] - conj X[L-k]) W_2L^-k,
           z = DFT^-1_L(Z) (unnormalized), y_2m = Re z_m, y_2m+1 = Im z_m.
           Im X[0] and Im X[L] are ignored, as the r2cb codelets do. */
        const cx<R> *in = reinterpret_cast<const cx<R> *>(a.in) + bin;
        const R *inr = reinterpret_cast<const R *>(a.in) + bin;
        R *outr = reinterpret_cast<R *>(a.out) + bout;
        const cx<R> *wk = reinterpret_cast<const cx<R> *>(a.aux);
        const cx<R> *rt = reinterpret_cast<const cx<R> *>(a.rtw);
        /* fused real-to-real kind; a compile-time FK_RM_NONE in the plain
           r2c / c2r instantiations */
        const unsigned rm = (IO == FK_IO_R2C_RR || IO == FK_IO_C2R_RR) ? a.rmode : 0u;
        constexpr long long N2 = 2LL * L;
        /* bin k of the fused real-to-real kinds: HC2R reads (r_k, i_k) from
           the halfcomplex line; REDFT01 forms exp(+i pi k / 2n)(x_k - i x_{n-k})
           (x_n = 0), RODFT01 the same on the reversed line */
        auto rbin = [&](int k) -> cx<R> {
            if (rm == FK_RM_HC2R)
                return mk(inr[(long long)k * a.is],
                          (k == 0 || k == L) ? (R)0 : inr[(N2 - k) * a.is]);
            const bool s = rm == FK_RM_DST3;
            const long long q0 = s ? N2 - 1 - k : k, q1 = s ? k - 1 : N2 - k;
            const R xk = inr[q0 * a.is];
            const R xn = k == 0 ? (R)0 : inr[q1 * a.is];
            const cx<R> w = rt[2 * k];
            return mk(w.x * xk - w.y * xn, -(w.x * xn + w.y * xk));
        };
        /* bins X[k], X[L-k] (k <= L/2) are read together and give Z[k] and
           Z[L-k] with one twiddle (W_2L^(L-k) = -conj W_2L^k); they are
           written to LDS before the stages, which read Z as it is */
        constexpr int HK = L / 2;
        constexpr int NK2 = (HK + 1 + TPL - 1) / TPL;
        constexpr int GK = 8;
        /* swapped Z (the backward DFT runs as the forward one on swapped
           data): e = X[k] + conj X[L-k], d = X[k] - conj X[L-k],
           o = d conj(W_2L^k); Z[k] = e + i o and, by the symmetry above,
           Z[L-k] = conj e + i conj o */
        auto zpair = [&](cx<R> xk, cx<R> xc, cx<R> w, cx<R> &z0, cx<R> &z1) {
            const cx<R> e = mk(xk.x + xc.x, xk.y - xc.y);
            const cx<R> d = mk(xk.x - xc.x, xk.y + xc.y);
            const cx<R> o = mk(d.x * w.x + d.y * w.y, d.y * w.x - d.x * w.y);
            z0 = mk(e.y + o.x, e.x - o.y);
            z1 = mk(o.x - e.y, e.x + o.y);
        };
        if constexpr (IO == FK_IO_C2R_XR) {
            /* last stage of the inverse column DFT over the tile's NL input
               rows (row p = the line's batch index i0, n1 = i1; the planner
               sets nb[0] = NL, so a tile is one n1): for every bin k
               u_n2 = sum_p W_NL^(-n2*p) W_n0^(-n1*p) Y_p[k], n0 = a.tN,
               which is line n2's half spectrum (its own output row) */
            constexpr long long NONE = -0x7fffffffffffffffLL - 1; /* INT64_MIN */
            if (u == 0 && real) lbase[lt] = act ? bin : NONE;
            const long long n1 = tile % a.nb[1];
            cx<R> wp[NL];
            sfor<0, NL>([&](auto pc) {
                constexpr int pp = decltype(pc)::value;
                const cx<R> t = pp == 0 ? mk((R)1, (R)0) : tw2<R>(a, n1 * pp);
                wp[pp] = mk(t.x, -t.y);
            });
            __syncthreads();
            const cx<R> *inb = reinterpret_cast<const cx<R> *>(a.in);
            /* bin k of every row of the tile: inverse DFT through the
               forward one on swapped data */
            auto coldft = [&](int k, cx<R> *v) {
                sfor<0, NL>([&](auto pc) {
                    constexpr int pp = decltype(pc)::value;
                    const long long ib = lbase[pp];
                    const cx<R> y = ib != NONE ? inb[ib + (long long)k * a.is] : mk((R)0, (R)0);
                    v[pp] = swp(pp == 0 ? y : mul(y, wp[pp]));
                });
                Dft<R, NL, 1>::run(v);
            };
            for (int k = tid; k <= HK; k += NT) {
                const cx<R> w = wk[k];
                cx<R> v[NL], vm[NL];
                coldft(k, v);
                if (2 * k != L) coldft(L - k, vm);
                sfor<0, NL>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    cx<R> xk = swp(v[q]);
                    cx<R> xc = 2 * k == L ? xk : swp(vm[q]);
                    if (k == 0) {
                        xk.y = 0;
                        xc.y = 0;
                    }
                    cx<R> z0, z1;
                    zpair(xk, xc, w, z0, z1);
                    lds[G::idx(q, k)] = z0;
                    if (k != 0 && 2 * k != L) lds[G::idx(q, L - k)] = z1;
                });
            }
            __syncthreads();
        } else {
            /* loads in groups of 8 (compile-time count) so that each group
               is in flight at once, then its LDS stores */
            sfor<0, (NK2 + GK - 1) / GK>([&](auto gc) {
                constexpr int g = decltype(gc)::value;
                constexpr int cnt = (NK2 - g * GK) < GK ? (NK2 - g * GK) : GK;
                cx<R> xa[cnt], xb[cnt], w[cnt];
                sfor<0, cnt>([&](auto ic) {
                    constexpr int i = decltype(ic)::value;
                    const int k = u + (g * GK + i) * TPL;
                    if (k <= HK) {
                        xa[i] = !act ? mk((R)0, (R)0)
                                : rm == FK_RM_NONE ? in[(long long)k * a.is] : rbin(k);
                        xb[i] = !act ? mk((R)0, (R)0)
                                : rm == FK_RM_NONE ? in[(long long)(L - k) * a.is] : rbin(L - k);
                        w[i] = wk[k];
                    }
                });
                sfor<0, cnt>([&](auto ic) {
                    constexpr int i = decltype(ic)::value;
                    const int k = u + (g * GK + i) * TPL;
                    if (k <= HK) {
                        cx<R> xk = xa[i], xc = xb[i];
                        if (k == 0) {
                            xk.y = 0;
                            xc.y = 0;
                        }
                        cx<R> z0, z1;
                        zpair(xk, xc, w[i], z0, z1);
                        lds_at(k) = z0;
                        if (k != 0 && 2 * k != L) lds_at(L - k) = z1;
                    }
                });
            });
            SY::sync();
        }
        auto gst = [&](int k, cx<R> v) {
            /* v = swap(z_k) */
            v.x *= sc; v.y *= sc;
            if (!act) return;
            if (rm == FK_RM_DCT3 || rm == FK_RM_DST3) {
                /* y_2i = v_i (i < L), y_{2i+1} = v_{n-1-i}: v_m goes to 2m
                   below L, to 2(n-1-m)+1 above; RODFT01 negates odd outputs */
                const long long m0 = 2LL * k, m1 = 2LL * k + 1;
                const long long d0 = m0 < L ? 2 * m0 : 2 * (N2 - 1 - m0) + 1;
                const long long d1 = m1 < L ? 2 * m1 : 2 * (N2 - 1 - m1) + 1;
                R y0 = v.y, y1 = v.x;
                if (rm == FK_RM_DST3) {
                    if (d0 & 1) y0 = -y0;
                    if (d1 & 1) y1 = -y1;
                }
                outr[d0 * a.os] = y0;
                outr[d1 * a.os] = y1;
            } else if (fl & FK_REAL_UNIT) {
                reinterpret_cast<cx<R> *>(outr)[k] = mk(v.y, v.x);
            } else {
                outr[(2 * k) * a.os] = v.y;
                outr[(2 * k + 1) * a.os] = v.x;
            }
        };
        auto zld = [&](auto j) -> cx<R> { return lds_at(j); };
        run_stages<R, RL, 0, L, TPL, true, false, SY>(u, tw, lds_at, zld, gst);
    } else if constexpr (IO == FK_IO_BLUE) {
        /* Bluestein (dft/bluestein.c:82-128 conventions): a_j = x_j c_j,
           c_j = exp(-i pi j^2/n); conv = IDFT_L(DFT_L(a) * B),
           B = DFT_L(conj c)/L; y_k = c_k conv_k, k < n. */
        const cx<R> *in = reinterpret_cast<const cx<R> *>(a.in) + bin;
        cx<R> *out = reinterpret_cast<cx<R> *>(a.out);
        const cx<R> *chirp = reinterpret_cast<const cx<R> *>(a.aux);
        const cx<R> *Bt = reinterpret_cast<const cx<R> *>(a.aux2);
        const int n = (int)a.nvalid;
        auto gld = [&](int j) -> cx<R> {
            if (!act || j >= n) return mk((R)0, (R)0);
            cx<R> v = in[j * a.is];
            if (swi) v = swp(v);
            return mul(v, chirp[j]);
        };
        auto lst = [&](auto k, cx<R> v) { lds_at(k) = v; };
        run_stages<R, RL, 0, L, TPL, false, true, SY>(u, tw, lds_at, gld, lst);
        auto mid = [&](int j) -> cx<R> { return swp(mul(lds_at(j), Bt[j])); };
        auto gst = [&](int k, cx<R> v) {
            if (k < n && act) {
                cx<R> y = mul(swp(v), chirp[k]);
                y.x *= sc; y.y *= sc;
                out[bout + k * a.os] = swo ? swp(y) : y;
            }
        };
        run_stages<R, RL, 0, L, TPL, true, false, SY>(u, tw, lds_at, mid, gst);
    }
}

template <typename R, int NL, int MODE, int IO, int PAD, class RL, int TPL>
__global__ void __launch_bounds__(TPL * NL)
stockham(fk_args a)
{
    extern __shared__ unsigned char smem_raw[];
    /* row mapping with whole lines inside one wave: wave-level exchange */
    constexpr bool wave_lines = MODE == FK_ROW && TPL <= 64 && 64 % TPL == 0;
    using SY = typename std::conditional<wave_lines, SyncWave, SyncAll>::type;
    stockham_body<R, NL, MODE, IO, PAD, RL, TPL, TPL * NL, SY>(a, remap_tile(a.xcd_remap),
                                                               smem_raw);
}

template <class T>
struct AtOff;
template <int OFF, int ALIGN>
struct AtOff<At<OFF, ALIGN>> {
    static constexpr int v = OFF;
};

/*
 * Row-mapped plain pass with one stage-0 butterfly per thread (L / R0 =
 * TPL) whose workgroups each run several tiles: the grid holds the
 * workgroups the device keeps resident, a workgroup steps through tiles
 * blockIdx.x + k*gridDim.x, and the next tile's stage-0 inputs are loaded
 * into registers while the current tile runs its stages, so HBM reads
 * overlap the LDS stages that otherwise follow each tile's loads.  For
 * long lines (one or a few per workgroup) the whole batch is only a couple
 * of workgroup rounds, which stockham<> runs with load, compute and store
 * phases in lock-step.  Same arithmetic as stockham<>: identical output.
 */
template <typename R, int NL, int PAD, class RL, int TPL>
__global__ void __launch_bounds__(TPL * NL, 4)
stockham_pipe(fk_args a)
{
    constexpr int L = RL::L(), R0 = RL::r[0], NB0 = L / R0;
    static_assert(NB0 == TPL, "one stage-0 butterfly per thread");
    using G = Geo<R, L, TPL, NL, FK_ROW, PAD>;
    constexpr bool wave_lines = TPL <= 64 && 64 % TPL == 0;
    using SY = typename std::conditional<wave_lines, SyncWave, SyncAll>::type;
    extern __shared__ unsigned char smem_raw[];
    cx<R> *lds = reinterpret_cast<cx<R> *>(smem_raw);
    const int tid = threadIdx.x, lt = tid / TPL, u = tid % TPL;
    const long long ntiles = (a.nlines + NL - 1) / NL, step = gridDim.x;
    const unsigned fl = a.flags;
    const bool swi = fl & FK_SWAP_IN, swo = fl & FK_SWAP_OUT;
    const R sc = (fl & FK_SCALE) ? (R)a.scale : (R)1;
    const long long is = a.is, os = a.os;
    const cx<R> *tw = reinterpret_cast<const cx<R> *>(a.tw);
    const cx<R> *in = reinterpret_cast<const cx<R> *>(a.in);
    cx<R> *out = reinterpret_cast<cx<R> *>(a.out);
    /* line of tile t for this thread: input and output offsets, computed
       once per tile when its inputs are fetched; lanes past the last line
       use line 0 (in bounds) and store nothing */
    auto line_of = [&](long long t, long long &bin, long long &bout) -> bool {
        const long long line = t * NL + lt;
        const bool act = line < a.nlines;
        long long q = act ? line : 0;
        const long long i0 = q % a.nb[0];
        q /= a.nb[0];
        const long long i1 = q % a.nb[1], i2 = q / a.nb[1];
        bin = i0 * a.ids[0] + i1 * a.ids[1] + i2 * a.ids[2];
        bout = i0 * a.ods[0] + i1 * a.ods[1] + i2 * a.ods[2];
        return act;
    };
    long long t = blockIdx.x;
    if (t >= ntiles) return;
    cx<R> cur[R0], nxt[R0];
    /* stage-0 inputs of the thread's butterfly */
    auto fetch = [&](long long bin, cx<R> *v) {
        const cx<R> *p = in + bin + u * is;
        sfor<0, R0>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            const cx<R> x = p[(long long)(r * NB0) * is];
            v[r] = swi ? swp(x) : x;
        });
    };
    long long bin, bout;
    bool act = line_of(t, bin, bout);
    fetch(bin, cur);
    auto lds_at = [&](auto j) -> cx<R> & { return lds[G::idx(lt, j)]; };
    /* output post-op of stockham<> without the four-step twiddle (the
       planner runs passes that need it with stockham<>): scale, swap */
    auto post = [&](cx<R> v) -> cx<R> {
        v.x *= sc;
        v.y *= sc;
        return swo ? swp(v) : v;
    };
    for (;;) {
        const long long tn = t + step;
        const bool more = tn < ntiles;
        long long bin_n = 0, bout_n = 0;
        bool act_n = false;
        if (more) {
            act_n = line_of(tn, bin_n, bout_n);
            fetch(bin_n, nxt);
        }
        cx<R> *outu = out + bout + u * os;
        auto first = [&](auto j) -> cx<R> {
            constexpr int r = AtOff<typename std::decay<decltype(j)>::type>::v / NB0;
            return cur[r];
        };
        auto gst = [&](auto k, cx<R> v) {
            if (NL == 1 || act) outu[(long long)((int)k - u) * os] = post(v);
        };
        run_stages<R, RL, 0, L, TPL, false, false, SY>(u, tw, lds_at, first, gst);
        if (!more) break;
        sfor<0, R0>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            cur[r] = nxt[r];
        });
        t = tn;
        bout = bout_n;
        act = act_n;
        /* the next tile's stage-0 LDS writes follow every read of this one */
        __syncthreads();
    }
}

/*
 * Row-mapped plain pass whose LDS exchange moves the real parts, then the
 * imaginary parts, through one array of reals: half the LDS per line of
 * stockham<> (more workgroups resident per CU to hide HBM latency) for
 * four workgroup barriers per exchange instead of two.  Stage s computes in
 * registers, writes its outputs' real parts, every thread reads the real
 * parts of its next-stage inputs, then the same for the imaginary parts.
 */
template <class RL, int S, int L, int TPL>
struct XsGeo {
    static constexpr int Rs = RL::r[S];
    static constexpr int Ns = RL::ns(S);
    static constexpr int NB = L / Rs;
    static constexpr int BPT = (NB + TPL - 1) / TPL;
};

template <typename R, class RL, int S, int L, int TPL, class PosF, class LastF>
__device__ __forceinline__ void run_xs(int u, const cx<R> *__restrict__ tw, R *buf, PosF &&idx,
                                       cx<R> (&v)[XsGeo<RL, S, L, TPL>::BPT][XsGeo<RL, S, L, TPL>::Rs],
                                       LastF &&last)
{
    using X = XsGeo<RL, S, L, TPL>;
    constexpr int Rs = X::Rs, Ns = X::Ns, NB = X::NB, BPT = X::BPT;
    sfor<0, BPT>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const int b = u + i * TPL;
        if ((NB % TPL == 0) || b < NB) {
            if constexpr (S > 0) {
                const int k = b % Ns;
                const cx<R> *w = tw + RL::twoff(S) + k * (Rs - 1);
                sfor<1, Rs>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    v[i][r] = mul(v[i][r], w[r - 1]);
                });
            }
            Dft<R, Rs, 1>::run(v[i]);
        }
    });
    auto each_out = [&](auto &&f) {
        sfor<0, BPT>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            const int b = u + i * TPL;
            if ((NB % TPL == 0) || b < NB) {
                const int k = b % Ns, base = (b - k) * Rs + k;
                sfor<0, Rs>([&](auto rc) {
                    constexpr int r = decltype(rc)::value;
                    f(base + r * Ns, v[i][r]);
                });
            }
        });
    };
    if constexpr (S == RL::n - 1) {
        each_out([&](int k, cx<R> x) { last(k, x); });
    } else {
        using Y = XsGeo<RL, S + 1, L, TPL>;
        cx<R> w[Y::BPT][Y::Rs];
        auto each_in = [&](auto &&f) {
            sfor<0, Y::BPT>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                const int b = u + i * TPL;
                if ((Y::NB % TPL == 0) || b < Y::NB) {
                    sfor<0, Y::Rs>([&](auto rc) {
                        constexpr int r = decltype(rc)::value;
                        f(b + r * Y::NB, w[i][r]);
                    });
                }
            });
        };
        each_out([&](int k, cx<R> x) { buf[idx(k)] = x.x; });
        __syncthreads();
        each_in([&](int j, cx<R> &x) { x.x = buf[idx(j)]; });
        __syncthreads();
        each_out([&](int k, cx<R> x) { buf[idx(k)] = x.y; });
        __syncthreads();
        each_in([&](int j, cx<R> &x) { x.y = buf[idx(j)]; });
        __syncthreads();
        run_xs<R, RL, S + 1, L, TPL>(u, tw, buf, idx, w, last);
    }
}

template <typename R, int NL, int MODE, int PAD, class RL, int TPL>
__global__ void __launch_bounds__(TPL * NL)
stockham_xs(fk_args a)
{
    extern __shared__ unsigned char smem_raw[];
    constexpr int L = RL::L();
    constexpr int LP = L + L / PAD + 2;
    const long long tile = remap_tile(a.xcd_remap);
    /* row mapping: a line's threads are consecutive, its reals one padded
       LDS row; column mapping: consecutive threads are consecutive lines
       (unit-stride batch), LDS laid out [element][line] */
    const int tid = threadIdx.x;
    const int lt = MODE == FK_ROW ? tid / TPL : tid % NL;
    const int u = MODE == FK_ROW ? tid % TPL : tid / NL;
    const long long line = tile * NL + lt;
    const bool act = line < a.nlines;
    long long q = act ? line : 0;
    const long long i0 = q % a.nb[0];
    q /= a.nb[0];
    const long long i1 = q % a.nb[1], i2 = q / a.nb[1];
    const long long bin = i0 * a.ids[0] + i1 * a.ids[1] + i2 * a.ids[2];
    const long long bout = i0 * a.ods[0] + i1 * a.ods[1] + i2 * a.ods[2];
    const cx<R> *tw = reinterpret_cast<const cx<R> *>(a.tw);
    const cx<R> *in = reinterpret_cast<const cx<R> *>(a.in) + bin;
    cx<R> *out = reinterpret_cast<cx<R> *>(a.out);
    const unsigned fl = a.flags;
    const bool swi = fl & FK_SWAP_IN, swo = fl & FK_SWAP_OUT;
    /* output scale: multiplying by 1 when unscaled is exact and costs less
       than selecting between scaled and unscaled values */
    const R sc = (fl & FK_SCALE) ? (R)a.scale : (R)1;
    R *buf = reinterpret_cast<R *>(smem_raw) + (MODE == FK_ROW ? lt * LP : lt);
    auto idx = [](int j) { return MODE == FK_ROW ? j + j / PAD : j * NL; };
    /* output post-op as in stockham_body: four-step twiddle anchored in the
       FP64 table every ANCH outputs of a butterfly, scale, swap */
    constexpr int NS_LAST = RL::ns(RL::n - 1);
    constexpr int ANCH = sizeof(R) == 8 ? 4 : 16;
    cx<double> tw_cur = mk(1.0, 0.0), tw_step = mk(1.0, 0.0);
    if (fl & FK_TW_OUT) tw_step = tw2d(a, (long long)NS_LAST * i0);
    auto gst = [&](int k, cx<R> v) {
        if (fl & FK_TW_OUT) {
            if ((k / NS_LAST) % ANCH == 0) tw_cur = tw2d(a, (long long)k * i0);
            else tw_cur = mul(tw_cur, tw_step);
            v = mul(v, mk((R)tw_cur.x, (R)tw_cur.y));
        }
        v.x *= sc; v.y *= sc;
        if (act) out[bout + k * a.os] = swo ? swp(v) : v;
    };
    using X0 = XsGeo<RL, 0, L, TPL>;
    cx<R> v[X0::BPT][X0::Rs];
    sfor<0, X0::BPT>([&](auto ic) {
        constexpr int i = decltype(ic)::value;
        const int b = u + i * TPL;
        if ((X0::NB % TPL == 0) || b < X0::NB) {
            sfor<0, X0::Rs>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                cx<R> x = act ? in[(long long)(b + r * X0::NB) * a.is] : mk((R)0, (R)0);
                v[i][r] = swi ? swp(x) : x;
            });
        }
    });
    run_xs<R, RL, 0, L, TPL>(u, tw, buf, idx, v, gst);
}

template <typename R, int NL, int MODE, int PAD, class RL, int TPL>
constexpr unsigned stockham_xs_lds()
{
    return (unsigned)(NL * (MODE == FK_ROW ? RL::L() + RL::L() / PAD + 2 : RL::L()) * sizeof(R));
}

template <typename R, int NL, int MODE, int IO, int PAD, class RL, int TPL, int NT = TPL * NL>
constexpr unsigned stockham_lds()
{
    using G = Geo<R, RL::L(), TPL, NL, MODE, PAD>;
    return G::lds_elems() * sizeof(cx<R>) + ((NL + 1) & ~1) * sizeof(long long) +
           (NT > TPL * NL ? 64 * sizeof(cx<R>) : 0);
}

} // namespace fk
#endif
