"""Complex DFT parity: the HIP engine (through fftw_plan_*/fftw_execute of
libfftw3_amd.so) against the long-double oracle on the same seeded inputs.

Coverage mirrors tests/check.pl: every 1-D size 1..100 and powers of two
(:185-196), 2-D sizes, random ranks and strides (:222-274), in/out of
place x forward/backward (:149-168), plus the configuration sizes (4096 x
1024, 1000, 1001, 2187, prime 4093, 65536, 2^20) and both precisions.
Bars: relative L2 <= 1e-13 (FP64) and <= 2e-6 (FP32).
"""
import numpy as np
import pytest

import fftw3_amd as fa
import oracle
from gpu_util import TOL, c2c_many, guru_c2c, mask_for

pytestmark = pytest.mark.gpu

SMALL = list(range(1, 101))
POW2 = [128, 256, 512, 1024, 2048, 4096, 8192, 16384]
ODD = [243, 625, 729, 1000, 1001, 2187, 3125, 4093, 4095, 6000, 8191]


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("n", SMALL)
def test_small_sizes(eng, prec, n):
    for sign in (-1, 1):
        y, ref, _ = c2c_many(prec, [n], 3, sign, seed=n)
        assert oracle.rel_l2(y, ref) <= TOL[prec], (n, sign)


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("n", POW2 + ODD)
def test_medium_sizes(eng, prec, n):
    for sign in (-1, 1):
        y, ref, _ = c2c_many(prec, [n], 2, sign, seed=n + 7)
        err = oracle.rel_l2(y, ref)
        assert err <= TOL[prec], (n, sign, err)


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("n", [64, 1000, 4096, 65536])
def test_inplace(eng, prec, n):
    y, ref, _ = c2c_many(prec, [n], 4, -1, seed=3, inplace=True)
    assert oracle.rel_l2(y, ref) <= TOL[prec]


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("n", [65536, 1 << 18, 1 << 20, 3 * (1 << 16), 65537])
def test_large_1d(eng, prec, n):
    y, ref, _ = c2c_many(prec, [n], 1, -1, seed=11)
    err = oracle.rel_l2(y, ref)
    assert err <= TOL[prec], err


def test_config_4096x1024_fp64(eng):
    """1-D c2c FP64, N=4096, howmany=1024, out of place, seed 1."""
    y, ref, _ = c2c_many("double", [4096], 1024, -1, seed=1)
    err = oracle.rel_l2(y, ref)
    assert err <= 1e-13, err


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("shape", [(4, 6), (16, 16), (30, 30), (7, 128), (64, 64), (32, 32, 32),
                                   (2, 3, 4, 5), (5, 1, 7), (128, 1000)])
def test_multid(eng, prec, shape):
    for sign in (-1, 1):
        y, ref, _ = c2c_many(prec, list(shape), 2, sign, seed=len(shape))
        assert oracle.rel_l2(y, ref) <= TOL[prec], (shape, sign)


@pytest.mark.parametrize("prec", ["double", "float"])
def test_guru_strided(eng, prec):
    # n=12 elements at stride 3, 5 batches at distance 40; output stride 2
    dims, hdims = [(12, 3, 2)], [(5, 40, 31)]
    y, ref = guru_c2c(prec, dims, hdims, -1, 400, 400)
    m = mask_for(dims, hdims, 400)
    assert oracle.rel_l2(y[m], ref[m]) <= TOL[prec]


@pytest.mark.parametrize("prec", ["double", "float"])
def test_guru_interleaved_batch(eng, prec):
    # batch-contiguous layout: element j of transform h at h + j*H (column FFTs)
    n, H = 256, 48
    dims, hdims = [(n, H, H)], [(H, 1, 1)]
    y, ref = guru_c2c(prec, dims, hdims, 1, n * H, n * H)
    assert oracle.rel_l2(y, ref) <= TOL[prec]


@pytest.mark.parametrize("prec", ["double", "float"])
def test_guru_rank2_howmany2(eng, prec):
    dims = [(6, 50, 1), (8, 5, 6)]
    hdims = [(3, 300, 48), (2, 1, 144)]
    y, ref = guru_c2c(prec, dims, hdims, -1, 900, 288)
    m = mask_for(dims, hdims, 288)
    assert oracle.rel_l2(y[m], ref[m]) <= TOL[prec]


def test_guru64_and_inplace_strided(eng):
    dims, hdims = [(100, 2, 2)], [(7, 203, 203)]
    y, ref = guru_c2c("double", dims, hdims, -1, 1500, 1500, inplace=True)
    assert oracle.rel_l2(y, ref) <= 1e-13


def test_rank0_and_howmany0(eng):
    x = np.arange(20, dtype=np.complex128) * (1 + 2j)
    bi = fa.DeviceBuffer.from_array(x)
    bo = fa.DeviceBuffer(x.nbytes)
    p = fa.plan_guru_dft("double", [], [(10, 2, 1)], bi.ptr, bo.ptr, -1)
    p.execute()
    assert np.array_equal(bo.download(np.complex128, 10), x[::2])
    p = fa.plan_many_dft("double", [8], 0, bi.ptr, None, 1, 8, bo.ptr, None, 1, 8, -1)
    p.execute()


def test_invalid_arguments_return_null(eng):
    lib = fa.lib()
    b = fa.DeviceBuffer(1024)
    n = (fa.ctypes.c_int * 1)(0)
    assert not lib.fftw_plan_many_dft(1, n, 1, b.ptr, None, 1, 1, b.ptr, None, 1, 1, -1, fa.FFTW_ESTIMATE)
    n[0] = 8
    assert not lib.fftw_plan_many_dft(1, n, -1, b.ptr, None, 1, 1, b.ptr, None, 1, 1, -1, fa.FFTW_ESTIMATE)
    # in-place with mismatched strides is unsolvable (dft/problem.c:95-99)
    assert not lib.fftw_plan_many_dft(1, n, 2, b.ptr, None, 1, 8, b.ptr, None, 2, 16, -1, fa.FFTW_ESTIMATE)


def test_new_array_execute_and_copy_plan(eng):
    n, H = 1000, 3
    x = oracle.uniform_complex(n * H, 5)
    a = fa.DeviceBuffer.from_array(x)
    b = fa.DeviceBuffer(x.nbytes)
    c = fa.DeviceBuffer.from_array(x[::-1].copy())
    d = fa.DeviceBuffer(x.nbytes)
    p = fa.plan_many_dft("double", [n], H, a.ptr, None, 1, n, b.ptr, None, 1, n, -1)
    q = fa.lib().fftw_copy_plan(p.h)
    assert q == p.h
    p.execute_dft(c.ptr, d.ptr)
    ref = oracle.fftn(x[::-1].reshape(H, n), -1, axes_batch=1).reshape(-1)
    assert oracle.rel_l2(d.download(np.complex128), ref) <= 1e-13
    fa.lib().fftw_destroy_plan(q)  # drops the extra reference only
    p.execute()
    ref2 = oracle.fftn(x.reshape(H, n), -1, axes_batch=1).reshape(-1)
    assert oracle.rel_l2(b.download(np.complex128), ref2) <= 1e-13


def test_input_preserved_out_of_place(eng):
    for n in (4096, 65536, 4093):
        x = oracle.uniform_complex(n, 9)
        a = fa.DeviceBuffer.from_array(x)
        b = fa.DeviceBuffer(x.nbytes)
        p = fa.plan_many_dft("double", [n], 1, a.ptr, None, 1, n, b.ptr, None, 1, n, -1)
        p.execute()
        assert np.array_equal(a.download(np.complex128), x)
