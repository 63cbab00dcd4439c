"""The benchmark's own plans, at full size, against the oracle.

The benchmark plans with FFTW_MEASURE, which times the catalog's variants
and keeps the fastest, so the timed plans are not the ESTIMATE plans the
other tests use.  Each problem here is planned exactly as the benchmark
plans it and its forward output is compared with the oracle: sampled bins of
the full transform (long-double sums with exact index reduction) where the
whole transform is too large for the oracle, whole batches where it is not.
The inverse transforms are pinned by the round trip against a forward
transform that the oracle has pinned (c2r(r2c x) = N x for every x fixes
c2r on all Hermitian input once r2c is right).
"""
import numpy as np
import pytest

import fftw3_amd as fa
import oracle
from gpu_util import TOL

pytestmark = pytest.mark.gpu


def bins_err(got, want):
    return np.sqrt(np.sum(np.abs(got - want) ** 2) / np.sum(np.abs(want) ** 2))


def sample_bins(shape, count, seed):
    rng = np.random.default_rng(seed)
    corners = [[0] * len(shape), [1] * len(shape), [n // 2 for n in shape], [n - 1 for n in shape]]
    rand = np.stack([rng.integers(0, n, count) for n in shape], axis=1)
    return np.concatenate([np.array(corners), rand])


def test_measure_c2c3d_512_fp64(eng):
    """3-D c2c FP64 512^3 in place, forward and backward planned with
    FFTW_MEASURE as the benchmark does; forward bins and backward bins (of
    the forward output) against the oracle, and the round trip."""
    n = 512
    N = n ** 3
    a = fa.DeviceBuffer(N * 16)
    pf = fa.plan_many_dft("double", [n] * 3, 1, a.ptr, None, 1, N, a.ptr, None, 1, N, -1,
                          fa.FFTW_MEASURE)
    pb = fa.plan_many_dft("double", [n] * 3, 1, a.ptr, None, 1, N, a.ptr, None, 1, N, +1,
                          fa.FFTW_MEASURE)
    print(pf.sprint(), pb.sprint(), flush=True)
    a.fill_uniform("double", 2 * N, 2, 0)
    x = a.download(np.complex128)
    pf.execute()
    X = a.download(np.complex128)
    bins = sample_bins((n, n, n), 8, 11)
    want = oracle.dft_bins_nd(x.reshape(n, n, n), bins, -1)
    Xr = X.reshape(n, n, n)
    got = np.array([Xr[tuple(b)] for b in bins])
    assert bins_err(got, want) <= TOL["double"], bins_err(got, want)
    pb.execute()
    y = a.download(np.complex128)
    wantb = oracle.dft_bins_nd(Xr, bins, +1)
    yr = y.reshape(n, n, n)
    gotb = np.array([yr[tuple(b)] for b in bins])
    assert bins_err(gotb, wantb) <= TOL["double"], bins_err(gotb, wantb)
    assert oracle.rel_l2(y / N, x) <= TOL["double"]
    pf.destroy()
    pb.destroy()


def test_measure_r2c_c2r_2048sq_x64_fp32(eng):
    """2-D r2c / c2r FP32 2048^2 x 64 out of place, FFTW_MEASURE: the first
    and last transforms of the r2c against the oracle's whole 2-D r2c, every
    transform's c2r(r2c x) = N x."""
    n, H = 2048, 64
    N, M = n * n, n * (n // 2 + 1)
    a = fa.DeviceBuffer(H * N * 4)
    c = fa.DeviceBuffer(H * M * 8)
    r = fa.DeviceBuffer(H * N * 4)
    pr = fa.plan_many_dft_r2c("float", [n, n], H, a.ptr, None, 1, N, c.ptr, None, 1, M,
                              fa.FFTW_MEASURE)
    pc = fa.plan_many_dft_c2r("float", [n, n], H, c.ptr, None, 1, M, r.ptr, None, 1, N,
                              fa.FFTW_MEASURE)
    print(pr.sprint(), pc.sprint(), flush=True)
    a.fill_uniform("float", H * N, 2, 1)
    pr.execute()
    for h in (0, H - 1):
        x = a.download(np.float32, N, offset=h * N * 4).astype(np.float64).reshape(n, n)
        X = c.download(np.complex64, M, offset=h * M * 8).astype(np.complex128)
        ref = oracle.rfftn(x).reshape(-1)
        assert oracle.rel_l2(X, ref) <= TOL["float"], h
    pc.execute()
    x = a.download(np.float32).astype(np.float64)
    y = r.download(np.float32).astype(np.float64)
    assert oracle.rel_l2(y / N, x) <= TOL["float"]
    pr.destroy()
    pc.destroy()


def test_measure_c2c_4096x1024_fp64(eng):
    """1-D c2c FP64 4096 x 1024 out of place, FFTW_MEASURE: every transform
    against the oracle."""
    n, H = 4096, 1024
    a = fa.DeviceBuffer(n * H * 16)
    b = fa.DeviceBuffer(n * H * 16)
    p = fa.plan_many_dft("double", [n], H, a.ptr, None, 1, n, b.ptr, None, 1, n, -1,
                         fa.FFTW_MEASURE)
    print(p.sprint(), flush=True)
    a.fill_uniform("double", 2 * n * H, 1, 0)
    p.execute()
    x = a.download(np.complex128).reshape(H, n)
    y = b.download(np.complex128)
    ref = oracle.fftn(x, -1, axes_batch=1).reshape(-1)
    assert oracle.rel_l2(y, ref) <= TOL["double"]
    p.destroy()


def test_measure_redft10_2048sq_x16_fp32(eng):
    """2-D DCT-II FP32 2048^2 x 16, FFTW_MEASURE: the first and last
    transforms against the oracle."""
    n, H = 2048, 16
    N = n * n
    a = fa.DeviceBuffer(H * N * 4)
    b = fa.DeviceBuffer(H * N * 4)
    k = [fa.FFTW_REDFT10, fa.FFTW_REDFT10]
    p = fa.plan_many_r2r("float", [n, n], H, a.ptr, None, 1, N, b.ptr, None, 1, N, k,
                         fa.FFTW_MEASURE)
    print(p.sprint(), flush=True)
    a.fill_uniform("float", H * N, 7, 0)
    p.execute()
    for h in (0, H - 1):
        x = a.download(np.float32, N, offset=h * N * 4).astype(np.float64).reshape(n, n)
        y = b.download(np.float32, N, offset=h * N * 4).astype(np.float64)
        ref = oracle.r2rn(x, k).reshape(-1)
        assert oracle.rel_l2(y, ref) <= TOL["float"], h
    p.destroy()


@pytest.mark.parametrize("variant", ["fp64", "fp32", "mfma"])
def test_measure_2p24_x8_sampled_bins(eng, variant):
    """c2c N = 2^24, howmany = 8, N(0,1) input (seed 4), planned with
    FFTW_MEASURE (the matrix-core leg with FFTW_AMD_MFMA): sampled bins of
    three transforms against the oracle's direct long-double sums."""
    n, H = 1 << 24, 8
    prec = "double" if variant == "fp64" else "float"
    cs = 16 if prec == "double" else 8
    flags = fa.FFTW_MEASURE | (fa.FFTW_AMD_MFMA if variant == "mfma" else 0)
    a = fa.DeviceBuffer(n * H * cs)
    b = fa.DeviceBuffer(n * H * cs)
    p = fa.plan_many_dft(prec, [n], H, a.ptr, None, 1, n, b.ptr, None, 1, n, -1, flags)
    print(p.sprint(), flush=True)
    if variant == "mfma":
        assert all(s.startswith("mfma16") for s in p.pass_info()), p.pass_info()
    a.fill_gauss(prec, 2 * n * H, 4, 0)
    p.execute()
    dt = np.complex128 if prec == "double" else np.complex64
    rng = np.random.default_rng(9)
    for h in (0, 4, 7):
        x = a.download(dt, n, offset=h * n * cs).astype(np.complex128)
        y = b.download(dt, n, offset=h * n * cs).astype(np.complex128)
        ks = np.concatenate([[0, 1, n // 2, n - 1], rng.integers(0, n, 12)])
        want = oracle.dft_bins(x, ks, -1)
        assert bins_err(y[ks], want) <= TOL[prec], (h, bins_err(y[ks], want))
    p.destroy()
