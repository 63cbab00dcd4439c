"""Real-to-real kinds on the GPU (plan_many_r2r / plan_guru{,64}_r2r /
plan_r2r_{1,2,3}d, execute_r2r) against the oracle on the same inputs.

Covers every kind of api/fftw3.h:89-100 at small, odd, prime, power-of-two
and large sizes, both precisions, in-place and out-of-place, batches,
multi-dimensional mixed kinds, guru strides (transposed and padded), new
arrays at execute, the size-1 and size-2 special cases, input
preservation, the NULL cases (size-1 REDFT00, unknown kind) and wisdom.
Bars: relative L2 of TOL (1e-13 FP64, 2e-6 FP32) against the long-double
oracle."""
import ctypes

import numpy as np
import pytest

import fftw3_amd as fa
import oracle
from gpu_util import TOL, inputs_r, rdt

pytestmark = pytest.mark.gpu

ALL = list(range(11))
NAMES = ["r2hc", "hc2r", "dht", "e00", "e01", "e10", "e11", "o00", "o01", "o10", "o11"]
SIZES = [1, 2, 3, 4, 5, 8, 15, 16, 17, 64, 100, 127, 128, 1000, 1024, 4093, 4096, 65536]


def many(prec, kind, n, howmany, seed, inplace=False, flags=fa.FFTW_ESTIMATE):
    x = inputs_r(n * howmany, seed, prec)
    a = fa.DeviceBuffer.from_array(x)
    o = a if inplace else fa.DeviceBuffer(x.nbytes)
    p = fa.plan_many_r2r(prec, [n], howmany, a.ptr, None, 1, n, o.ptr, None, 1, n, [kind], flags)
    if not (flags & fa.FFTW_ESTIMATE):
        a.upload(x)
    p.execute()
    y = o.download(rdt(prec))
    xin = None if inplace else a.download(rdt(prec))
    p.destroy()
    return x, y, xin


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("kind", ALL, ids=NAMES)
def test_r2r_1d_sizes(prec, kind):
    for n in SIZES:
        if kind == 3 and n == 1:
            continue
        h = 3 if n <= 4096 else 2
        print("r2r", prec, NAMES[kind], n, flush=True)
        x, y, xin = many(prec, kind, n, h, 100 + n)
        ref = oracle.r2rn(x.astype(np.float64).reshape(h, n), [kind], axes_batch=1).reshape(-1)
        assert oracle.rel_l2(y, ref) < TOL[prec], (NAMES[kind], n)
        assert np.array_equal(xin, x), "input not preserved"


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("kind,sizes", [(3, [3, 9, 65, 129, 1025, 4097]), (7, [1, 7, 63, 255, 1023, 4095])],
                         ids=["e00", "o00"])
def test_r2r_symmetric_extension_sizes(prec, kind, sizes):
    """REDFT00 / RODFT00 sizes whose logical transform 2(n-1) / 2(n+1) is
    twice a power of two: the extension is read inside the real kernel."""
    for n in sizes:
        x, y, xin = many(prec, kind, n, 4, 300 + n)
        ref = oracle.r2rn(x.astype(np.float64).reshape(4, n), [kind], axes_batch=1).reshape(-1)
        assert oracle.rel_l2(y, ref) < TOL[prec], (NAMES[kind], n)
        assert np.array_equal(xin, x)
    x, y, _ = many(prec, kind, sizes[-1], 3, 7, inplace=True)
    ref = oracle.r2rn(x.astype(np.float64).reshape(3, -1), [kind], axes_batch=1).reshape(-1)
    assert oracle.rel_l2(y, ref) < TOL[prec]


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("kind", ALL, ids=NAMES)
def test_r2r_inplace(prec, kind):
    for n in (6, 7, 256, 1001):
        x, y, _ = many(prec, kind, n, 5, 7 + n, inplace=True)
        ref = oracle.r2rn(x.astype(np.float64).reshape(5, n), [kind], axes_batch=1).reshape(-1)
        assert oracle.rel_l2(y, ref) < TOL[prec], (NAMES[kind], n)


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("kinds,shape", [((5, 5), (64, 48)),
                                         ((4, 10), (33, 20)),
                                         ((0, 2, 7), (6, 10, 12)),
                                         ((3, 1, 6), (9, 8, 5)),
                                         ((6, 8), (1, 40)),
                                         ((6, 10), (64, 32)),
                                         ((10, 6, 6), (16, 12, 8)),
                                         ((5, 5, 5), (32, 32, 32))])
@pytest.mark.parametrize("inplace", [False, True])
def test_r2r_multid(prec, kinds, shape, inplace):
    N = int(np.prod(shape))
    x = inputs_r(2 * N, 21, prec)
    a = fa.DeviceBuffer.from_array(x)
    o = a if inplace else fa.DeviceBuffer(x.nbytes)
    p = fa.plan_many_r2r(prec, list(shape), 2, a.ptr, None, 1, N, o.ptr, None, 1, N, list(kinds))
    p.execute()
    y = o.download(rdt(prec))
    ref = oracle.r2rn(x.astype(np.float64).reshape((2,) + shape), kinds, axes_batch=1).reshape(-1)
    assert oracle.rel_l2(y, ref) < TOL[prec]
    p.destroy()


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("wide", [False, True])
def test_r2r_guru_strides(prec, wide):
    """Transposed input, padded output rows, two howmany dims."""
    n0, n1, h0, h1 = 12, 20, 3, 2
    kinds = [5, 9]
    span_in = n0 * n1 * h0 * h1
    pad = n1 + 5
    span_out = n0 * pad * h0 * h1
    x = inputs_r(span_in, 31, prec)
    # input: element (i0, i1, b0, b1) at i1*n0 + i0 + b0*(n0*n1) + b1*(n0*n1*h0)
    dims = [(n0, 1, pad), (n1, n0, 1)]
    hdims = [(h0, n0 * n1, n0 * pad), (h1, n0 * n1 * h0, n0 * pad * h0)]
    a = fa.DeviceBuffer.from_array(x)
    o = fa.DeviceBuffer(span_out * x.itemsize)
    o.upload(np.full(span_out, -7, rdt(prec)))
    p = fa.plan_guru_r2r(prec, dims, hdims, a.ptr, o.ptr, kinds, wide=wide)
    p.execute()
    y = o.download(rdt(prec))
    out = np.full(span_out, -7.0)
    oracle.guru_r2r(dims, hdims, kinds, x.astype(np.float64), out)
    assert oracle.rel_l2(y, out) < TOL[prec]
    # padding untouched
    assert np.all(y.reshape(h1, h0, n0, pad)[..., n1:] == -7)
    p.destroy()


@pytest.mark.parametrize("prec", ["double", "float"])
def test_r2r_execute_new_arrays_and_wrappers(prec):
    L = fa.lib()
    P = fa.prefix(prec)
    n0, n1, n2 = 8, 6, 10
    x = inputs_r(n0 * n1 * n2, 41, prec)
    a, b = fa.DeviceBuffer.from_array(x), fa.DeviceBuffer(x.nbytes)
    a2, b2 = fa.DeviceBuffer.from_array(x[::-1].copy()), fa.DeviceBuffer(x.nbytes)
    h1 = getattr(L, P + "plan_r2r_1d")(n0 * n1 * n2, a.ptr, b.ptr, 6, fa.FFTW_ESTIMATE)
    h2 = getattr(L, P + "plan_r2r_2d")(n0 * n1, n2, a.ptr, b.ptr, 4, 2, fa.FFTW_ESTIMATE)
    h3 = getattr(L, P + "plan_r2r_3d")(n0, n1, n2, a.ptr, b.ptr, 5, 8, 0, fa.FFTW_ESTIMATE)
    cases = [(fa.Plan(h1, prec), [6], (n0 * n1 * n2,)), (fa.Plan(h2, prec), [4, 2], (n0 * n1, n2)),
             (fa.Plan(h3, prec), [5, 8, 0], (n0, n1, n2))]
    for p, kinds, shape in cases:
        p.execute()
        ref = oracle.r2rn(x.astype(np.float64).reshape(shape), kinds)
        assert oracle.rel_l2(b.download(rdt(prec)), ref.reshape(-1)) < TOL[prec]
        p.execute_r2r(a2.ptr, b2.ptr)
        ref2 = oracle.r2rn(x[::-1].astype(np.float64).reshape(shape), kinds)
        assert oracle.rel_l2(b2.download(rdt(prec)), ref2.reshape(-1)) < TOL[prec]
        p.destroy()


@pytest.mark.parametrize("prec", ["double", "float"])
def test_r2r_null_cases(prec):
    b = fa.DeviceBuffer(1 << 12)
    with pytest.raises(RuntimeError):
        fa.plan_many_r2r(prec, [1], 1, b.ptr, None, 1, 1, b.ptr, None, 1, 1, [3])
    with pytest.raises(RuntimeError):
        fa.plan_many_r2r(prec, [8], 1, b.ptr, None, 1, 8, b.ptr, None, 1, 8, [11])
    # size-1 REDFT00 inside a multi-D problem is also undefined
    with pytest.raises(RuntimeError):
        fa.plan_many_r2r(prec, [4, 1], 1, b.ptr, None, 1, 4, b.ptr, None, 1, 4, [5, 3])
    # a rank-0 / all-trivial problem is a copy
    x = inputs_r(16, 3, prec)
    a, o = fa.DeviceBuffer.from_array(x), fa.DeviceBuffer(x.nbytes)
    p = fa.plan_many_r2r(prec, [1], 16, a.ptr, None, 1, 1, o.ptr, None, 1, 1, [0])
    p.execute()
    assert np.array_equal(o.download(rdt(prec)), x)
    p.destroy()


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("kind", [5, 4, 6, 10, 0])
def test_r2r_measure_and_wisdom(prec, kind):
    L = fa.lib()
    P = fa.prefix(prec)
    getattr(L, P + "forget_wisdom")()
    n, h = 4096, 64
    x, y, _ = many(prec, kind, n, h, 55, flags=fa.FFTW_MEASURE)
    ref = oracle.r2rn(x.astype(np.float64).reshape(h, n), [kind], axes_batch=1).reshape(-1)
    assert oracle.rel_l2(y, ref) < TOL[prec]
    # the timed plan is now wisdom: WISDOM_ONLY planning succeeds
    x2, y2, _ = many(prec, kind, n, h, 56, flags=fa.FFTW_ESTIMATE | fa.FFTW_WISDOM_ONLY)
    ref2 = oracle.r2rn(x2.astype(np.float64).reshape(h, n), [kind], axes_batch=1).reshape(-1)
    assert oracle.rel_l2(y2, ref2) < TOL[prec]


@pytest.mark.parametrize("prec,shape,kinds", [("float", (1024, 1024), (5, 5)),
                                              ("double", (512, 512), (4, 4)),
                                              ("float", (1 << 20,), (6,))])
def test_r2r_large(prec, shape, kinds):
    N = int(np.prod(shape))
    x = inputs_r(N, 61, prec)
    a, o = fa.DeviceBuffer.from_array(x), fa.DeviceBuffer(x.nbytes)
    p = fa.plan_many_r2r(prec, list(shape), 1, a.ptr, None, 1, N, o.ptr, None, 1, N, list(kinds))
    p.execute()
    ref = oracle.r2rn(x.astype(np.float64).reshape(shape), kinds).reshape(-1)
    assert oracle.rel_l2(o.download(rdt(prec)), ref) < TOL[prec]
    p.destroy()


R2R_SIZES1 = list(range(1, 21)) + [31, 32, 63, 64, 65, 100, 127, 128, 255, 256, 257, 1000, 1024, 2049]
R2R_SIZES2 = list(range(1, 17)) + [32, 33, 64]
R2R_SIZES3 = list(range(1, 10)) + [16]


def _layout(rng, shape):
    """Random physical layout (permuted, padded, one dim reversed) of a
    logical array: per-dim strides, offset of index 0, buffer span."""
    nd = len(shape)
    order = list(rng.permutation(nd)) if rng.random() < 0.4 else list(range(nd))
    strides = [0] * nd
    s = 1
    for d in reversed(order):
        pad = int(rng.choice([1, 1, 1, 2, 3]))
        strides[d] = s * pad
        s *= shape[d] * pad
        if rng.random() < 0.2:
            s += int(rng.integers(1, 5))
    base = 0
    if rng.random() < 0.25:
        d = int(rng.integers(0, nd))
        base += (shape[d] - 1) * strides[d]
        strides[d] = -strides[d]
    lo = base + sum(min(0, (n - 1) * st) for n, st in zip(shape, strides))
    hi = base + sum(max(0, (n - 1) * st) for n, st in zip(shape, strides))
    return strides, base - lo, hi - lo + 1


def _offsets(shape, strides, base):
    idx = np.full(1, base, np.int64)
    for n, st in zip(shape, strides):
        idx = (idx[:, None] + st * np.arange(n, dtype=np.int64)[None, :]).reshape(-1)
    return idx


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("block", range(10))
def test_random_r2r_guru_sweep(prec, block):
    """Seeded sweep in the manner of tests/check.pl over r2r problems: rank
    1-3, random kinds per dim, howmany 1-3, permuted / padded / reversed
    (negative-stride) layouts, in place and out of place, ESTIMATE and
    MEASURE.  Positions the transform must not write stay untouched."""
    rng = np.random.default_rng(5000 + 31 * block + (prec == "float"))
    rs = np.dtype(rdt(prec)).itemsize
    for case in range(15):
        rank = int(rng.choice([1, 1, 2, 2, 3]))
        pool = {1: R2R_SIZES1, 2: R2R_SIZES2, 3: R2R_SIZES3}[rank]
        dims = [int(rng.choice(pool)) for _ in range(rank)]
        kinds = [int(rng.integers(0, 11)) for _ in range(rank)]
        kinds = [5 if (k == 3 and n == 1) else k for k, n in zip(kinds, dims)]
        H = int(rng.integers(1, 4))
        shape = [H] + dims
        inplace = rng.random() < 0.3
        si, bi, span_i = _layout(rng, shape)
        so, bo, span_o = (si, bi, span_i) if inplace else _layout(rng, shape)
        measure = rng.random() < 0.2
        N = int(np.prod(shape))
        x = oracle.uniform(N, 9000 + case, block).astype(rdt(prec))
        buf_i = np.zeros(span_i, rdt(prec))
        buf_i[_offsets(shape, si, bi)] = x
        dI = fa.DeviceBuffer.from_array(buf_i)
        dO = dI if inplace else fa.DeviceBuffer.from_array(np.zeros(span_o, rdt(prec)))
        gd = [(shape[k], si[k], so[k]) for k in range(1, len(shape))]
        hd = [(shape[0], si[0], so[0])]
        what = (shape, kinds, si, so, bi, bo, inplace, measure)
        print("r2r sweep", prec, block, case, what, flush=True)
        p = fa.plan_guru_r2r(prec, gd, hd, dI.ptr + bi * rs, dO.ptr + bo * rs, kinds,
                             fa.FFTW_MEASURE if measure else fa.FFTW_ESTIMATE, wide=bool(case % 2))
        if measure:
            dI.upload(buf_i)
            if not inplace:
                dO.upload(np.zeros(span_o, rdt(prec)))
        p.execute()
        y = dO.download(rdt(prec))
        p.destroy()
        ref = oracle.r2rn(x.astype(np.float64).reshape(shape), kinds, axes_batch=1).reshape(-1)
        oo = _offsets(shape, so, bo)
        assert oracle.rel_l2(y[oo], ref) <= TOL[prec], what
        if not inplace:
            m = np.ones(span_o, bool)
            m[oo] = False
            assert not np.any(y[m]), what
            assert np.array_equal(dI.download(rdt(prec)), buf_i), ("input changed", what)


@pytest.mark.parametrize("prec", ["double", "float"])
def test_r2r_flops_print_wisdom_copy(prec):
    """fftw_flops counts the plan's own arithmetic (same order as the
    2.5 N log2 N of libbench2/mflops.c), print_plan names the kinds, MEASURE wisdom survives export / forget /
    import and makes WISDOM_ONLY planning succeed, copy_plan refcounts."""
    L = fa.lib()
    P = fa.prefix(prec)
    n0, n1, h = 64, 48, 4
    N = n0 * n1
    x = inputs_r(N * h, 71, prec)
    a, o = fa.DeviceBuffer.from_array(x), fa.DeviceBuffer(x.nbytes)
    kinds = [fa.FFTW_REDFT10, fa.FFTW_RODFT11]
    getattr(L, P + "forget_wisdom")()
    p = fa.plan_many_r2r(prec, [n0, n1], h, a.ptr, None, 1, N, o.ptr, None, 1, N, kinds, fa.FFTW_MEASURE)
    add, mul, fma = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    getattr(L, P + "flops")(p.h, ctypes.byref(add), ctypes.byref(mul), ctypes.byref(fma))
    nominal = 2.5 * N * np.log2(N) * h
    total = add.value + mul.value + 2 * fma.value
    assert 0.5 * nominal <= total <= 3 * nominal, (total, nominal)
    desc = p.sprint()
    assert "r2r" in desc and "e10" in desc and "o11" in desc
    q = fa.Plan(getattr(L, P + "copy_plan")(p.h), prec)
    p.destroy()
    a.upload(x)
    q.execute()
    ref = oracle.r2rn(x.astype(np.float64).reshape(h, n0, n1), kinds, axes_batch=1).reshape(-1)
    assert oracle.rel_l2(o.download(rdt(prec)), ref) < TOL[prec]
    q.destroy()
    ptr = L.fftw_export_wisdom_to_string()
    s = ctypes.cast(ptr, ctypes.c_char_p).value
    L._libc.free(ptr)
    assert s.count(b"(#x") >= 1
    getattr(L, P + "forget_wisdom")()
    with pytest.raises(RuntimeError):
        fa.plan_many_r2r(prec, [n0, n1], h, a.ptr, None, 1, N, o.ptr, None, 1, N, kinds,
                         fa.FFTW_ESTIMATE | fa.FFTW_WISDOM_ONLY)
    assert L.fftw_import_wisdom_from_string(s) == 1
    w = fa.plan_many_r2r(prec, [n0, n1], h, a.ptr, None, 1, N, o.ptr, None, 1, N, kinds,
                         fa.FFTW_ESTIMATE | fa.FFTW_WISDOM_ONLY)
    a.upload(x)
    w.execute()
    assert oracle.rel_l2(o.download(rdt(prec)), ref) < TOL[prec]
    w.destroy()


@pytest.mark.parametrize("prec", ["double", "float"])
def test_r2r_edge_shapes(prec):
    """howmany = 0 (nothing written), rank 3 with two howmany dims (more
    batch dims than one launch addresses), a length with a large prime
    factor (8193 = 3 * 2731) and a 2^18 line (half-length real transform
    through a four-step)."""
    b = fa.DeviceBuffer(4096)
    b.upload(np.full(1024, 5, np.float32))
    p = fa.plan_guru_r2r(prec, [(8, 1, 1)], [(0, 8, 8)], b.ptr, b.ptr, [5])
    p.execute()
    assert np.all(b.download(np.float32) == 5)
    p.destroy()
    shape = (2, 3, 4, 5, 6)          # (h1, h0, n0, n1, n2)
    N = int(np.prod(shape))
    x = inputs_r(N, 81, prec)
    a, o = fa.DeviceBuffer.from_array(x), fa.DeviceBuffer(x.nbytes)
    dims = [(4, 30, 30), (5, 6, 6), (6, 1, 1)]
    hd = [(3, 120, 120), (2, 360, 360)]
    kinds = [7, 0, 6]
    p = fa.plan_guru_r2r(prec, dims, hd, a.ptr, o.ptr, kinds)
    p.execute()
    ref = oracle.r2rn(x.astype(np.float64).reshape(shape), kinds, axes_batch=2).reshape(-1)
    assert oracle.rel_l2(o.download(rdt(prec)), ref) < TOL[prec]
    p.destroy()
    for kind, n, h in ((5, 8193, 2), (4, 8193, 2), (10, 8193, 2), (5, 1 << 18, 2), (1, 1 << 18, 2)):
        x, y, xin = many(prec, kind, n, h, 90 + kind)
        ref = oracle.r2rn(x.astype(np.float64).reshape(h, n), [kind], axes_batch=1).reshape(-1)
        assert oracle.rel_l2(y, ref) < TOL[prec], (NAMES[kind], n)
        assert np.array_equal(xin, x)


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("kind", [0, 2, 5, 9, 1, 4])
def test_r2r_output_offset_by_one_real(prec, kind):
    """Regression: kinds whose kernels read the input in pairs or quads
    (R2HC, DHT, REDFT10, RODFT10) or write the output in pairs (HC2R,
    REDFT01) must execute when the other array is only aligned to one real
    (the alignment requirement is per array, as at planning)."""
    n, h = 4096, 8
    rs = np.dtype(rdt(prec)).itemsize
    x = inputs_r(n * h, 111 + kind, prec)
    a = fa.DeviceBuffer(x.nbytes + 16 * rs)
    o = fa.DeviceBuffer(x.nbytes + 16 * rs)
    a.upload(np.concatenate([x, np.zeros(16, rdt(prec))]))
    o.upload(np.zeros(n * h + 16, rdt(prec)))
    pi, po = a.ptr, o.ptr + rs           # output starts one real in
    if kind in (1, 4):                   # real-output kernels: offset the input instead
        pi, po = a.ptr + rs, o.ptr
        a.upload(np.concatenate([np.zeros(1, rdt(prec)), x, np.zeros(15, rdt(prec))]))
    p = fa.plan_many_r2r(prec, [n], h, pi, None, 1, n, po, None, 1, n, [kind])
    p.execute()
    y = o.download(rdt(prec))
    y = y[1:1 + n * h] if kind not in (1, 4) else y[:n * h]
    ref = oracle.r2rn(x.astype(np.float64).reshape(h, n), [kind], axes_batch=1).reshape(-1)
    assert oracle.rel_l2(y, ref) < TOL[prec]
    p.destroy()
