"""fftw_flops / fftw_estimate_cost from the plan (api/flops.c:23-40,
kernel/ops.c:24-60).

The counts are those of the kernels the plan runs.  The CPU tests pin the
per-butterfly counts against derivations written out by hand below; the GPU
tests pin whole plans: one radix-16 pass (n = 16) and the four radix-8
stages of the 4096-point kernel, each from the stage structure of a
Stockham pass (L/R butterflies per stage, R-1 twiddle products on every
stage after the first).

Convention: a product feeding an addition is one fma, a complex product is
2 mul + 2 fma; the fftw_flops total is add + mul + 2 fma.
"""
import ctypes

import pytest

import fftw3_amd as fa


def dft_ops(n):
    L = fa.lib()
    f = L.fk_dft_ops
    f.restype = None
    f.argtypes = [ctypes.c_int] + [ctypes.POINTER(ctypes.c_double)] * 3
    a, m, x = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    f(n, ctypes.byref(a), ctypes.byref(m), ctypes.byref(x))
    return a.value, m.value, x.value


# Hand derivations of one forward butterfly (complex values: one complex
# add = 2 real adds).
#  2: y0 = a + b, y1 = a - b                                   -> 4 add
#  4: a = x0+x2, b = x0-x2, c = x1+x3, d = x1-x3 (8 add);
#     y0 = a+c, y2 = a-c, y1 = b - i d, y3 = b + i d (8 add)   -> 16 add
#  8 = 4 x 2: two 4-point DFTs (32 add); twiddles W8^1, W8^2, W8^3 on the
#     second column: W8^2 = -i is free, W8^1 and W8^3 cost 2 add + 2 mul
#     each (multiply by 1/sqrt2 after a sum); four 2-point DFTs (16 add)
#                                                   -> 52 add, 4 mul
#  16 = 4 x 4: eight 4-point DFTs (128 add); nine twiddles W16^e,
#     e = k1*n2 for k1, n2 in 1..3 = {1,2,3,2,4,6,3,6,9}: e = 4 free;
#     e = 2, 6 (odd multiples of pi/4): 2 add + 2 mul, four of them;
#     e = 1, 3, 3, 9: general complex product 2 mul + 2 fma, four of them
#                                                   -> 136 add, 16 mul, 8 fma
#  odd prime p, H = (p-1)/2 pairs: s_j = x_j + x_{p-j}, d_j = x_j - x_{p-j}
#     (4H add), y0 = x0 + sum s_j (2H add), each output pair k accumulates
#     4 real sums of H products (4H fma) and combines them (4 add)
#                                                   -> 10H add, 4H^2 fma
HAND = {
    2: (4, 0, 0),
    4: (16, 0, 0),
    8: (52, 4, 0),
    16: (136, 16, 8),
    3: (10, 0, 4),
    5: (20, 0, 16),
    7: (30, 0, 36),
    11: (50, 0, 100),
    13: (60, 0, 144),
}


@pytest.mark.parametrize("n", sorted(HAND))
def test_butterfly_counts(n):
    assert dft_ops(n) == HAND[n]


def test_power_of_two_butterflies_beat_radix2():
    """A radix-2 decimation of 2^k points costs 5 N log2 N real flops minus
    the free twiddles; the specialised 4-, 8- and 16-point butterflies cost
    at most that."""
    for n, k in ((4, 2), (8, 3), (16, 4)):
        a, m, f = dft_ops(n)
        assert a + m + 2 * f <= 5 * n * k


def test_composites_are_sums_of_their_parts():
    """A x B butterflies: B A-point DFTs, A B-point DFTs, plus the
    inter-factor twiddles, none dearer than a complex product (2 mul + 2 fma,
    6 flops)."""
    for n, A, B in ((6, 2, 3), (10, 2, 5), (12, 4, 3), (9, 3, 3), (32, 8, 4)):
        a, m, f = dft_ops(n)
        a1, m1, f1 = dft_ops(A)
        a2, m2, f2 = dft_ops(B)
        base = (B * a1 + A * a2, B * m1 + A * m2, B * f1 + A * f2)
        tw = (A - 1) * (B - 1)
        assert a >= base[0] and m >= base[1] and f >= base[2]
        assert (a - base[0]) + (m - base[1]) + 2 * (f - base[2]) <= 6 * tw


def plan_flops(p):
    add, mul, fma = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    fa.fn(p.prec, "flops")(p.h, ctypes.byref(add), ctypes.byref(mul), ctypes.byref(fma))
    return add.value, mul.value, fma.value


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["double", "float"])
def test_plan_flops_n16(eng, prec):
    """n = 16: one radix-16 stage per transform, no twiddles."""
    H = 7
    a = fa.DeviceBuffer(16 * H * 16)
    b = fa.DeviceBuffer(16 * H * 16)
    p = fa.plan_many_dft(prec, [16], H, a.ptr, None, 1, 16, b.ptr, None, 1, 16, fa.FFTW_FORWARD,
                         fa.FFTW_ESTIMATE)
    info = p.pass_info()
    assert len(info) == 1 and "rad=16 " in info[0], info
    assert plan_flops(p) == (136 * H, 16 * H, 8 * H)
    # backward is the same arithmetic (re/im swaps are free)
    q = fa.plan_many_dft(prec, [16], H, a.ptr, None, 1, 16, b.ptr, None, 1, 16, fa.FFTW_BACKWARD,
                         fa.FFTW_ESTIMATE)
    assert plan_flops(q) == plan_flops(p)
    p.destroy()
    q.destroy()


@pytest.mark.gpu
def test_plan_flops_n4096(eng):
    """n = 4096 = 8*8*8*8 (the 4096-point row kernel): 4 stages of 512
    radix-8 butterflies (52 add + 4 mul each); stages 1..3 multiply 7 inputs
    per butterfly by a twiddle (2 mul + 2 fma each)."""
    H = 3
    a = fa.DeviceBuffer(16 * 4096 * H)
    b = fa.DeviceBuffer(16 * 4096 * H)
    p = fa.plan_many_dft("double", [4096], H, a.ptr, None, 1, 4096, b.ptr, None, 1, 4096,
                         fa.FFTW_FORWARD, fa.FFTW_ESTIMATE)
    info = p.pass_info()
    assert len(info) == 1 and "rad=8x8x8x8 " in info[0], info
    adds = 4 * 512 * 52
    muls = 4 * 512 * 4 + 3 * 512 * 7 * 2
    fmas = 3 * 512 * 7 * 2
    assert plan_flops(p) == (adds * H, muls * H, fmas * H)
    tot = adds + muls + 2 * fmas
    # the same order as the nominal 5 N log2 N of libbench2/mflops.c
    assert 0.5 < tot / (5 * 4096 * 12) < 1.0
    # estimate_cost: modelled seconds, at least the HBM time of the bytes
    est = fa.fn("double", "estimate_cost")(p.h)
    assert est >= 2 * 16 * 4096 * H / 8e12
    p.destroy()


@pytest.mark.gpu
def test_plan_flops_scale_with_passes(eng):
    """A two-pass split (n = 2^20 FP32) counts both passes plus the four-step
    twiddle; the count grows with howmany exactly linearly."""
    n = 1 << 20
    a = fa.DeviceBuffer(8 * n * 2)
    b = fa.DeviceBuffer(8 * n * 2)
    p1 = fa.plan_many_dft("float", [n], 1, a.ptr, None, 1, n, b.ptr, None, 1, n, fa.FFTW_FORWARD,
                          fa.FFTW_ESTIMATE)
    p2 = fa.plan_many_dft("float", [n], 2, a.ptr, None, 1, n, b.ptr, None, 1, n, fa.FFTW_FORWARD,
                          fa.FFTW_ESTIMATE)
    assert p1.npasses() >= 2
    f1, f2 = plan_flops(p1), plan_flops(p2)
    assert all(abs(y - 2 * x) <= 1e-9 * y for x, y in zip(f1, f2))
    tot = f1[0] + f1[1] + 2 * f1[2]
    assert 0.5 < tot / (5 * n * 20) < 1.5
    p1.destroy()
    p2.destroy()
