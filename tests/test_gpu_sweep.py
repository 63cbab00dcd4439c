"""Randomized and full-size parity checks on the GPU.

* A seeded sweep over problem shapes in the spirit of the reference's
  tests/check.pl + libbench2 problem generator (libbench2/problem.c:229-318):
  rank 1-3, sizes mixing powers of two, smooth composites and primes,
  howmany 1-4, permuted / padded / reversed (negative-stride) layouts,
  in-place and out-of-place, both signs and precisions.  Each result is
  compared with the long-double oracle on the logical arrays, and positions
  the transform must not write are checked untouched.
* One plan executed concurrently from several host threads on different
  arrays (only fftw_execute* is thread-safe, doc/threads.texi:233-246).
* The benchmark configurations at full size, through properties that do not
  need a full-size oracle: round trip backward(forward x) = N x (C2a 512^3
  FP64, C2b 2048^2 x 64 FP32 r2c/c2r), linearity, and exact sampled output
  bins from the oracle's direct DFT (C4, N = 2^24).
"""
import threading

import numpy as np
import pytest

import fftw3_amd as fa
import oracle
from gpu_util import TOL, cdt, rdt

pytestmark = pytest.mark.gpu

SIZES1 = list(range(1, 33)) + [60, 64, 97, 100, 128, 243, 256, 500, 512, 625, 1000, 1001, 1024,
                                2048, 2187, 4093, 4096, 8192, 10007, 65536]
SIZES2 = list(range(1, 25)) + [32, 48, 64, 100, 128]
SIZES3 = list(range(1, 13)) + [16, 32]


def layout(rng, shape):
    """Random physical layout of a logical array of `shape` (batch dim
    first): returns per-dim strides in elements (possibly negative), the
    element offset of logical index 0 and the span of the buffer."""
    nd = len(shape)
    order = list(rng.permutation(nd)) if rng.random() < 0.4 else list(range(nd))
    strides = [0] * nd
    s = 1
    for d in reversed(order):  # last in `order` is fastest
        pad = int(rng.choice([1, 1, 1, 2, 3]))
        strides[d] = s * pad
        s *= shape[d] * pad
        if rng.random() < 0.2:
            s += int(rng.integers(1, 5))
    base = 0
    if rng.random() < 0.25:  # one reversed dim: negative stride
        d = int(rng.integers(0, nd))
        base += (shape[d] - 1) * strides[d]
        strides[d] = -strides[d]
    lo = base + sum(min(0, (n - 1) * st) for n, st in zip(shape, strides))
    hi = base + sum(max(0, (n - 1) * st) for n, st in zip(shape, strides))
    return strides, base - lo, hi - lo + 1


def offsets(shape, strides, base):
    idx = np.full(1, base, np.int64)
    for n, st in zip(shape, strides):
        idx = (idx[:, None] + st * np.arange(n, dtype=np.int64)[None, :]).reshape(-1)
    return idx


def random_c2c(rng, prec):
    rank = int(rng.choice([1, 1, 1, 2, 2, 3]))
    pool = {1: SIZES1, 2: SIZES2, 3: SIZES3}[rank]
    dims = [int(rng.choice(pool)) for _ in range(rank)]
    H = int(rng.integers(1, 5))
    shape = [H] + dims
    inplace = rng.random() < 0.3
    si, bi, span_i = layout(rng, shape)
    if inplace:
        so, bo, span_o = si, bi, span_i
    else:
        so, bo, span_o = layout(rng, shape)
    sign = int(rng.choice([-1, 1]))
    return shape, si, bi, span_i, so, bo, span_o, inplace, sign


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("block", range(10))
def test_random_c2c_sweep(prec, block):
    rng = np.random.default_rng(1000 + 17 * block + (prec == "float"))
    cs = np.dtype(cdt(prec)).itemsize
    for case in range(15):
        shape, si, bi, span_i, so, bo, span_o, inplace, sign = random_c2c(rng, prec)
        measure = rng.random() < 0.2
        N = int(np.prod(shape))
        x = (oracle.uniform(2 * N, 7000 + case, block) ).astype(rdt(prec)).view(cdt(prec))
        buf_i = np.zeros(span_i, cdt(prec))
        oi = offsets(shape, si, bi)
        buf_i[oi] = x
        dI = fa.DeviceBuffer.from_array(buf_i)
        dO = dI if inplace else fa.DeviceBuffer.from_array(np.zeros(span_o, cdt(prec)))
        # guru dims: transform dims = shape[1:], howmany dim = shape[0]
        dims = [(shape[k], si[k], so[k]) for k in range(1, len(shape))]
        hd = [(shape[0], si[0], so[0])]
        p = fa.plan_guru_dft(prec, dims, hd, dI.ptr + bi * cs, dO.ptr + bo * cs, sign,
                             fa.FFTW_MEASURE if measure else fa.FFTW_ESTIMATE)
        if measure:  # MEASURE may overwrite both arrays
            dI.upload(buf_i)
            if not inplace:
                dO.upload(np.zeros(span_o, cdt(prec)))
        p.execute()
        y = dO.download(cdt(prec))
        p.destroy()
        ref = oracle.fftn(x.astype(np.complex128).reshape(shape), sign, axes_batch=1).reshape(-1)
        oo = offsets(shape, so, bo)
        err = oracle.rel_l2(y[oo], ref)
        what = (shape, si, so, bi, bo, inplace, sign)
        assert err <= TOL[prec], (what, err)
        if not inplace:
            m = np.ones(span_o, bool)
            m[oo] = False
            assert not np.any(y[m]), what


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("block", range(6))
def test_random_real_sweep(prec, block):
    """r2c then c2r on random real / half-complex layouts (permuted,
    padded, reversed), in and out of place by separate buffers."""
    rng = np.random.default_rng(5000 + 31 * block + (prec == "float"))
    rs = np.dtype(rdt(prec)).itemsize
    cs = 2 * rs
    for case in range(12):
        rank = int(rng.choice([1, 1, 2, 2, 3]))
        pool = {1: SIZES1, 2: SIZES2, 3: SIZES3}[rank]
        dims = [int(rng.choice(pool)) for _ in range(rank)]
        H = int(rng.integers(1, 4))
        rshape = [H] + dims
        cshape = [H] + dims[:-1] + [dims[-1] // 2 + 1]
        sr, br, span_r = layout(rng, rshape)
        sc, bc, span_c = layout(rng, cshape)
        x = oracle.uniform(int(np.prod(rshape)), 8000 + case, block).astype(rdt(prec))
        buf_r = np.zeros(span_r, rdt(prec))
        orr = offsets(rshape, sr, br)
        buf_r[orr] = x
        dR = fa.DeviceBuffer.from_array(buf_r)
        dC = fa.DeviceBuffer.from_array(np.zeros(span_c, cdt(prec)))
        d_f = [(rshape[k], sr[k], sc[k]) for k in range(1, rank + 1)]
        p = fa.plan_guru_dft_r2c(prec, d_f, [(H, sr[0], sc[0])], dR.ptr + br * rs, dC.ptr + bc * cs)
        p.execute()
        X = dC.download(cdt(prec))
        p.destroy()
        ref = oracle.rfftn(x.astype(np.float64).reshape(rshape), axes_batch=1).reshape(-1)
        oc = offsets(cshape, sc, bc)
        what = ("r2c", rshape, sr, sc, br, bc)
        assert oracle.rel_l2(X[oc], ref) <= TOL[prec], (what, oracle.rel_l2(X[oc], ref))
        m = np.ones(span_c, bool)
        m[oc] = False
        assert not np.any(X[m]), what
        # c2r of the spectrum into a fresh real layout
        so, bo, span_o = layout(rng, rshape)
        dO = fa.DeviceBuffer.from_array(np.zeros(span_o, rdt(prec)))
        d_b = [(rshape[k], sc[k], so[k]) for k in range(1, rank + 1)]
        q = fa.plan_guru_dft_c2r(prec, d_b, [(H, sc[0], so[0])], dC.ptr + bc * cs, dO.ptr + bo * rs)
        q.execute()
        y = dO.download(rdt(prec)).astype(np.float64)
        q.destroy()
        N = int(np.prod(dims))
        oo = offsets(rshape, so, bo)
        what = ("c2r", rshape, sc, so, bc, bo)
        assert oracle.rel_l2(y[oo] / N, x.astype(np.float64)) <= 2 * TOL[prec], what


@pytest.mark.parametrize("prec", ["double", "float"])
def test_negative_strides(prec):
    """Reversed input and output (stride -1 from the last element), the
    strided reversed batch of api/mktensor-iodims.h semantics."""
    n, H = 1000, 3
    cs = np.dtype(cdt(prec)).itemsize
    x = oracle.uniform(2 * n * H, 77).astype(rdt(prec)).view(cdt(prec))
    dI = fa.DeviceBuffer.from_array(x)
    dO = fa.DeviceBuffer(x.nbytes)
    # element j of batch h at (H-1-h)*n + (n-1-j): both dims reversed
    last = n * H - 1
    p = fa.plan_guru_dft(prec, [(n, -1, -1)], [(H, -n, -n)], dI.ptr + last * cs, dO.ptr + last * cs, -1)
    p.execute()
    y = dO.download(cdt(prec))
    xl = x.astype(np.complex128)[::-1].reshape(H, n)
    ref = oracle.fftn(xl, -1, axes_batch=1).reshape(-1)[::-1]
    assert oracle.rel_l2(y, ref) <= TOL[prec]
    p.destroy()


def test_concurrent_execute():
    """One plan, six host threads, different arrays, executed repeatedly."""
    n, H, T = 65536, 4, 6
    xs = [oracle.uniform_complex(n * H, 90 + t) for t in range(T)]
    ins = [fa.DeviceBuffer.from_array(x) for x in xs]
    outs = [fa.DeviceBuffer(xs[0].nbytes) for _ in range(T)]
    p = fa.plan_many_dft("double", [n], H, ins[0].ptr, None, 1, n, outs[0].ptr, None, 1, n, -1)
    errs = [None] * T

    def work(t):
        for _ in range(5):
            p.execute_dft(ins[t].ptr, outs[t].ptr)
        y = outs[t].download(np.complex128)
        errs[t] = oracle.rel_l2(y, oracle.fftn(xs[t].reshape(H, n), -1, axes_batch=1).reshape(-1))

    th = [threading.Thread(target=work, args=(t,)) for t in range(T)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert all(e is not None and e <= 1e-13 for e in errs), errs
    p.destroy()


def test_full_c2a_512cubed_roundtrip_and_linearity():
    """C2a: 3-D c2c FP64 512^3 in place: backward(forward x) = N x, and
    F(x + 2i y) = F(x) + 2i F(y) on the full array."""
    n = 512
    N = n ** 3
    a = fa.DeviceBuffer(N * 16)
    a.fill_uniform("double", 2 * N, 2, 0)
    x = a.download(np.complex128)
    pf = fa.plan_many_dft("double", [n] * 3, 1, a.ptr, None, 1, N, a.ptr, None, 1, N, -1)
    pb = fa.plan_many_dft("double", [n] * 3, 1, a.ptr, None, 1, N, a.ptr, None, 1, N, +1)
    pf.execute()
    X = a.download(np.complex128)
    pb.execute()
    y = a.download(np.complex128)
    assert oracle.rel_l2(y / N, x) <= 1e-13
    # linearity against a second input
    b = fa.DeviceBuffer(N * 16)
    b.fill_uniform("double", 2 * N, 2, 1)
    yv = b.download(np.complex128)
    pf.execute_dft(b.ptr, b.ptr)
    Y = b.download(np.complex128)
    a.upload(x + 2j * yv)
    pf.execute()
    Z = a.download(np.complex128)
    assert oracle.rel_l2(Z, X + 2j * Y) <= 1e-13
    # a few exact bins from the oracle's direct 3-D sum along one line
    # (separable: the 3-D bin (0, 0, k) is the 1-D DFT of the plane sums)
    plane = x.reshape(n, n, n).sum(axis=(0, 1))
    ks = np.array([0, 1, 7, 255, 511])
    want = oracle.dft_bins(plane, ks, -1)
    got = X.reshape(n, n, n)[0, 0, ks]
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) <= 1e-12
    pf.destroy()
    pb.destroy()


def test_full_c2b_r2c_c2r_roundtrip():
    """C2b: 2-D r2c / c2r FP32 2048^2 x 64 out of place: c2r(r2c(x)) = N x,
    and every transform's DC bin equals the sum of its input."""
    n, H = 2048, 64
    N, M = n * n, n * (n // 2 + 1)
    a = fa.DeviceBuffer(H * N * 4)
    a.fill_uniform("float", H * N, 2, 1)
    c = fa.DeviceBuffer(H * M * 8)
    r = fa.DeviceBuffer(H * N * 4)
    pr = fa.plan_many_dft_r2c("float", [n, n], H, a.ptr, None, 1, N, c.ptr, None, 1, M)
    pc = fa.plan_many_dft_c2r("float", [n, n], H, c.ptr, None, 1, M, r.ptr, None, 1, N)
    pr.execute()
    X = c.download(np.complex64).reshape(H, M)
    x = a.download(np.float32).reshape(H, N)
    dc = x.astype(np.float64).sum(axis=1)
    assert np.max(np.abs(X[:, 0].real - dc)) <= 2e-6 * np.sqrt(N) * 4
    pc.execute()
    y = r.download(np.float32).reshape(H, N)
    assert oracle.rel_l2(y.astype(np.float64) / N, x.astype(np.float64)) <= TOL["float"]
    pr.destroy()
    pc.destroy()


@pytest.mark.parametrize("prec", ["double", "float"])
def test_full_c4_2p24_sampled_bins(prec):
    """C4: c2c N = 2^24, howmany = 8: exact sampled bins of every transform
    from the oracle's direct long-double sum with exact index reduction."""
    n, H = 1 << 24, 8
    cs = np.dtype(cdt(prec)).itemsize
    a = fa.DeviceBuffer(n * H * cs)
    a.fill_gauss(prec, 2 * n * H, 4, 0)
    b = fa.DeviceBuffer(n * H * cs)
    p = fa.plan_many_dft(prec, [n], H, a.ptr, None, 1, n, b.ptr, None, 1, n, -1)
    p.execute()
    x = a.download(cdt(prec)).astype(np.complex128).reshape(H, n)
    y = b.download(cdt(prec)).astype(np.complex128).reshape(H, n)
    rng = np.random.default_rng(5)
    for h in (0, 3, 7):
        ks = np.concatenate([[0, 1, n // 2, n - 1], rng.integers(0, n, 4)])
        want = oracle.dft_bins(x[h], ks, -1)
        scale = np.sqrt(n)  # |Y_k| ~ sqrt(n) for N(0,1) inputs
        err = np.max(np.abs(y[h, ks] - want)) / scale
        assert err <= (1e-13 if prec == "double" else 2e-6) * 10, (h, err)
    p.destroy()


def test_beyond_2p31_elements():
    """64-bit indexing: 2^19 + 3 transforms of 4096 FP32 points (just over
    2^31 complex elements, 16 GiB); the first and the last transforms are
    checked against the oracle, and a middle one by its DC bin."""
    n, H = 4096, (1 << 19) + 3
    tot = n * H
    a = fa.DeviceBuffer(tot * 8)
    b = fa.DeviceBuffer(tot * 8)
    a.fill_uniform("float", 2 * tot, 6, 0)
    p = fa.plan_many_dft("float", [n], H, a.ptr, None, 1, n, b.ptr, None, 1, n, -1)
    p.execute()
    for h in (0, H // 2, H - 1):
        x = a.download(np.complex64, n, offset=h * n * 8).astype(np.complex128)
        y = b.download(np.complex64, n, offset=h * n * 8).astype(np.complex128)
        assert oracle.rel_l2(y, oracle.fftn(x, -1)) <= TOL["float"], h
    p.destroy()
    # the same arrays as 512 2-D problems of 2048^2 starting 12288 elements
    # in, so the last one ends past element 2^31
    sh = 12288
    m = 2048 * 2048
    q = fa.plan_many_dft("float", [2048, 2048], 512, a.ptr + sh * 8, None, 1, m, b.ptr + sh * 8, None, 1,
                         m, +1)
    q.execute()
    last = sh + 511 * m
    assert last + m > (1 << 31)
    x = a.download(np.complex64, m, offset=last * 8).astype(np.complex128).reshape(2048, 2048)
    y = b.download(np.complex64, m, offset=last * 8).astype(np.complex128).reshape(2048, 2048)
    assert oracle.rel_l2(y, oracle.fftn(x, +1)) <= TOL["float"]
    q.destroy()
