/*
 * Device-side building blocks of the Stockham engine (gfx950, wave64).
 *
 * What each butterfly computes is fixed by the DFT definition
 * (doc/reference.texi:1872-1902): forward DFT, sign -1, unnormalized.  The
 * backward transform is the forward one with real and imaginary parts
 * swapped on load and store, the identity FFTW's execute uses
 * (api/execute-dft.c:25-32, kernel/extract-reim.c:27).
 *
 * Butterflies are generated at compile time from templates:
 *   - odd primes p: the symmetric form over pairs (x_j, x_{p-j}) with
 *     constexpr cos/sin constants;
 *   - composite N = A*B: A-point DFTs on B columns, inter-factor twiddles
 *     W_N^{k1*n2} (multiples of pi/4 specialised), B-point DFTs on A rows,
 *     and a register permutation to natural order.
 * These replace FFTW's generated n1_* codelets (genfft/gen_notw.ml:58-160).
 */
#ifndef FK_DEVICE_H
#define FK_DEVICE_H

#include <hip/hip_runtime.h>
#include <type_traits>
#include "fk_args.h"

namespace fk {

template <typename R> struct cx { R x, y; };
template <> struct __align__(8) cx<float> { float x, y; };
template <> struct __align__(16) cx<double> { double x, y; };

template <typename R>
__device__ __forceinline__ cx<R> mk(R x, R y) { cx<R> r; r.x = x; r.y = y; return r; }
template <typename R>
__device__ __forceinline__ cx<R> add(cx<R> a, cx<R> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename R>
__device__ __forceinline__ cx<R> sub(cx<R> a, cx<R> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename R>
__device__ __forceinline__ cx<R> mul(cx<R> a, cx<R> b)
{
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R>
__device__ __forceinline__ cx<R> swp(cx<R> a) { return mk(a.y, a.x); }

/* ---- compile-time trigonometry (Taylor series on an octant) ---- */

constexpr double kPI = 3.14159265358979323846264338327950288;

constexpr double taylor_sin(double x)
{
    double t = x, s = x;
    for (int k = 1; k < 30; ++k) {
        t *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        s += t;
    }
    return s;
}

constexpr double taylor_cos(double x)
{
    double t = 1.0, s = 1.0;
    for (int k = 1; k < 30; ++k) {
        t *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        s += t;
    }
    return s;
}

/* cos and sin of 2*pi*m/n, exact octant folding on integers */
struct cs { double c, s; };
constexpr cs kcexp(long long m, long long n)
{
    m %= n;
    if (m < 0) m += n;
    long long N4 = 4 * n, M4 = 4 * m, q = n;
    unsigned oct = 0;
    if (M4 > N4 - M4) { M4 = N4 - M4; oct |= 4; }
    if (M4 - q > 0) { M4 -= q; oct |= 2; }
    if (M4 > q - M4) { M4 = q - M4; oct |= 1; }
    double th = 2.0 * kPI * (double)M4 / (double)N4;
    double c = taylor_cos(th), s = taylor_sin(th), t = 0;
    if (oct & 1) { t = c; c = s; s = t; }
    if (oct & 2) { t = c; c = -s; s = t; }
    if (oct & 4) { s = -s; }
    return cs{c, s};
}
constexpr double kcos(long long m, long long n) { return kcexp(m, n).c; }
constexpr double ksin(long long m, long long n) { return kcexp(m, n).s; }

constexpr bool is_prime(int n)
{
    if (n < 2) return false;
    for (int d = 2; d * d <= n; ++d)
        if (n % d == 0) return false;
    return true;
}

constexpr int first_factor(int n)
{
    if (n % 4 == 0 && n != 4) {
        /* powers of two: split into balanced power-of-two halves */
        int a = 4;
        while (a * a < n && n % (a * 2) == 0 && (n / (a * 2)) >= 2) a *= 2;
        if (n % a == 0) return a;
        return 4;
    }
    for (int d = 2; d * d <= n; ++d)
        if (n % d == 0) return d;
    return n;
}

template <int B, int E, class F>
__device__ __forceinline__ void sfor(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(f);
    }
}

/* v *= W_N^E = exp(-2*pi*i*E/N) */
template <typename R, int N, int E>
__device__ __forceinline__ void twc(cx<R> &v)
{
    constexpr int e = ((E % N) + N) % N;
    constexpr R h = (R)0.70710678118654752440084436210484903928;
    const R a = v.x, b = v.y;
    if constexpr (e == 0) {
    } else if constexpr (8 * e == N) {
        v.x = (a + b) * h; v.y = (b - a) * h;
    } else if constexpr (4 * e == N) {
        v.x = b; v.y = -a;
    } else if constexpr (8 * e == 3 * N) {
        v.x = (b - a) * h; v.y = -(a + b) * h;
    } else if constexpr (2 * e == N) {
        v.x = -a; v.y = -b;
    } else if constexpr (8 * e == 5 * N) {
        v.x = -(a + b) * h; v.y = (a - b) * h;
    } else if constexpr (4 * e == 3 * N) {
        v.x = -b; v.y = a;
    } else if constexpr (8 * e == 7 * N) {
        v.x = (a - b) * h; v.y = (a + b) * h;
    } else {
        constexpr R c = (R)kcos(e, N);
        constexpr R s = (R)-ksin(e, N);
        v.x = a * c - b * s;
        v.y = a * s + b * c;
    }
}

/* In-place forward DFT of v[0], v[S], ..., v[(N-1)*S] (compile-time N, S). */
template <typename R, int N, int S>
struct Dft {
    __device__ __forceinline__ static void run(cx<R> *v)
    {
        if constexpr (N == 1) {
        } else if constexpr (N == 2) {
            cx<R> a = v[0], b = v[S];
            v[0] = add(a, b);
            v[S] = sub(a, b);
        } else if constexpr (N == 4) {
            cx<R> a = add(v[0], v[2 * S]), b = sub(v[0], v[2 * S]);
            cx<R> c = add(v[S], v[3 * S]), d = sub(v[S], v[3 * S]);
            v[0] = add(a, c);
            v[2 * S] = sub(a, c);
            v[S] = mk(b.x + d.y, b.y - d.x);       /* b - i*d */
            v[3 * S] = mk(b.x - d.y, b.y + d.x);   /* b + i*d */
        } else if constexpr (is_prime(N)) {
            constexpr int H = (N - 1) / 2;
            cx<R> s[H + 1], d[H + 1];
            const cx<R> x0 = v[0];
            cx<R> y0 = x0;
            sfor<1, H + 1>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                s[j] = add(v[j * S], v[(N - j) * S]);
                d[j] = sub(v[j * S], v[(N - j) * S]);
                y0 = add(y0, s[j]);
            });
            sfor<1, H + 1>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                R ar = x0.x, ai = x0.y, br = 0, bi = 0;
                sfor<1, H + 1>([&](auto jc) {
                    constexpr int j = decltype(jc)::value;
                    constexpr R c = (R)kcos((long long)j * k, N);
                    constexpr R sn = (R)ksin((long long)j * k, N);
                    ar += c * s[j].x;
                    ai += c * s[j].y;
                    br += sn * d[j].x;
                    bi += sn * d[j].y;
                });
                v[k * S] = mk(ar + bi, ai - br);
                v[(N - k) * S] = mk(ar - bi, ai + br);
            });
            v[0] = y0;
        } else {
            constexpr int A = first_factor(N), B = N / A;
            /* A-point DFTs down the B columns */
            sfor<0, B>([&](auto n2c) {
                constexpr int n2 = decltype(n2c)::value;
                Dft<R, A, B * S>::run(v + n2 * S);
            });
            /* twiddles W_N^{k1*n2} */
            sfor<1, A>([&](auto k1c) {
                constexpr int k1 = decltype(k1c)::value;
                sfor<1, B>([&](auto n2c) {
                    constexpr int n2 = decltype(n2c)::value;
                    twc<R, N, k1 * n2>(v[(n2 + B * k1) * S]);
                });
            });
            /* B-point DFTs along the A rows */
            sfor<0, A>([&](auto k1c) {
                constexpr int k1 = decltype(k1c)::value;
                Dft<R, B, S>::run(v + B * k1 * S);
            });
            /* Y[k1 + A*k2] sits at v[(B*k1 + k2)*S]: permute */
            cx<R> t[N];
            sfor<0, A>([&](auto k1c) {
                constexpr int k1 = decltype(k1c)::value;
                sfor<0, B>([&](auto k2c) {
                    constexpr int k2 = decltype(k2c)::value;
                    t[k1 + A * k2] = v[(B * k1 + k2) * S];
                });
            });
            sfor<0, N>([&](auto ic) {
                constexpr int i = decltype(ic)::value;
                v[i * S] = t[i];
            });
        }
    }
};

/* ---- exact FP64 twiddles on the device ---- */

/* exp(-2*pi*i*m/n) in FP64, m reduced exactly and folded to an octant
   before sincospi (same reduction as kernel/trig.c:57-80). */
__device__ __forceinline__ void cexp_fp64(long long m, long long n, double &c,
                                          double &s)
{
    m %= n;
    if (m < 0) m += n;
    long long N4 = 4 * n, M4 = 4 * m, q = n;
    unsigned oct = 0;
    if (M4 > N4 - M4) { M4 = N4 - M4; oct |= 4; }
    if (M4 - q > 0) { M4 -= q; oct |= 2; }
    if (M4 > q - M4) { M4 = q - M4; oct |= 1; }
    double sn, cs_;
    sincospi((double)M4 / (double)(2 * n), &sn, &cs_);
    double t;
    if (oct & 1) { t = cs_; cs_ = sn; sn = t; }
    if (oct & 2) { t = cs_; cs_ = -sn; sn = t; }
    if (oct & 4) { sn = -sn; }
    c = cs_;
    s = -sn;
}

/* W_N^e in FP64 from the two-level table (see tw2) */
__device__ __forceinline__ cx<double> tw2d(const fk_args &a, long long e);

template <typename R>
__device__ __forceinline__ cx<R> tw2(const fk_args &a, long long e)
{
    const cx<double> w = tw2d(a, e);
    return mk((R)w.x, (R)w.y);
}

__device__ __forceinline__ cx<double> tw2d(const fk_args &a, long long e)
{
    /* W_N^e from the two-level FP64 table, W_N^e = lo[e % b] * hi[e / b].
       Four-step exponents are e = k*i0 < N; the reduction below only runs
       for exponents outside [0, N). */
    const long long N = a.tN, b = a.tb;
    if (e < 0 || e >= N) {
        e %= N;
        if (e < 0) e += N;
    }
    long long il, ih;
    if ((b & (b - 1)) == 0) {                  /* power-of-two split */
        il = e & (b - 1);
        ih = e >> __builtin_ctzll((unsigned long long)b);
    } else if (N <= 0x7fffffffLL) {            /* 32-bit division */
        const unsigned ue = (unsigned)e, ub = (unsigned)b, q = ue / ub;
        ih = q;
        il = ue - q * ub;
    } else {
        ih = e / b;
        il = e - ih * b;
    }
    const double *lo = a.tlo + 2 * il, *hi = a.thi + 2 * ih;
    return mk(lo[0] * hi[0] - lo[1] * hi[1], lo[0] * hi[1] + lo[1] * hi[0]);
}

} // namespace fk
#endif
