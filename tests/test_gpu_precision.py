"""Precision ladder and the split-precision matrix-core variant.

FP64 plans, FP32 plans (FP64-generated twiddles rounded to FP32) and FP32
plans whose radix-16 passes run as FP16 hi/lo matrix products with FP32
accumulation, all against the long-double oracle on Gaussian inputs.
Bars: rel-L2 <= 1e-13 (FP64) and <= 2e-6 (FP32 and the split variant).
"""
import numpy as np
import pytest

import fftw3_amd as fa
import oracle

pytestmark = pytest.mark.gpu


def test_mfma_layout(eng):
    """In-bounds probe of v_mfma_f32_32x32x8_f16 operand/result layouts:
    identity rows 0..7 times B gives B in rows 0..7 and zeros below."""
    buf = fa.DeviceBuffer(32 * 32 * 4)
    assert fa.lib().fftw_amd_mfma_probe(buf.ptr) == 0
    d = buf.download(np.float32).reshape(32, 32)
    want = np.zeros((32, 32), np.float32)
    want[:8] = (np.arange(8)[:, None] * 32 + np.arange(32)[None, :]) / 256.0
    assert np.array_equal(d, want)


def gauss_c(count, seed, dtype):
    return oracle.gauss(2 * count, seed).astype(dtype).view(
        np.complex64 if dtype == np.float32 else np.complex128)


@pytest.mark.parametrize("n,howmany", [(16, 64), (256, 16), (4096, 8), (65536, 4)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_mfma_variant_accuracy(eng, n, howmany, sign):
    x = gauss_c(n * howmany, 40 + n % 97, np.float32)
    a = fa.DeviceBuffer.from_array(x)
    b = fa.DeviceBuffer(x.nbytes)
    # FFTW_AMD_MFMA plans radix-16 passes only (and keys its own wisdom), so
    # the matrix-core kernel serves every pass whatever earlier plans timed
    p = fa.plan_many_dft("float", [n], howmany, a.ptr, None, 1, n, b.ptr, None, 1, n, sign,
                         fa.FFTW_ESTIMATE | fa.FFTW_AMD_MFMA)
    switched = fa.lib().fftw_amd_plan_use_mfma(p.h)
    assert switched >= 1
    rads = [s.split("rad=")[1].split()[0].split("x") for s in p.pass_info()]
    assert all(set(r) == {"16"} for r in rads), p.pass_info()
    p.execute()
    y = b.download(np.complex64)
    ref = oracle.fftn(x.astype(np.complex128).reshape(howmany, n), sign, axes_batch=1).reshape(-1)
    err = oracle.rel_l2(y, ref)
    print("mfma n=%d sign=%d rel_l2=%.3e" % (n, sign, err))
    assert err <= 2e-6, err


def test_mfma_planner_flag(eng):
    """FFTW_AMD_MFMA decomposes power-of-16 lengths into radix-16 passes that
    all run on the matrix-core kernel."""
    n, H = 1 << 20, 2
    x = gauss_c(n * H, 9, np.float32)
    a = fa.DeviceBuffer.from_array(x)
    b = fa.DeviceBuffer(x.nbytes)
    p = fa.plan_many_dft("float", [n], H, a.ptr, None, 1, n, b.ptr, None, 1, n, -1,
                         fa.FFTW_ESTIMATE | fa.FFTW_AMD_MFMA)
    info = p.pass_info()
    assert all("rad=16x16x16" in s or "rad=16x16 " in s or "rad=16 " in s for s in info), info
    p.execute()
    ref = oracle.fftn(x.astype(np.complex128).reshape(H, n), -1, axes_batch=1).reshape(-1)
    assert oracle.rel_l2(b.download(np.complex64), ref) <= 2e-6


def test_precision_ladder(eng):
    # 2^16 = 256 x 256: both passes have radix-16 stages only
    n, H = 1 << 16, 16
    x64 = gauss_c(n * H, 4, np.float64)
    ref = oracle.fftn(x64.reshape(H, n), -1, axes_batch=1).reshape(-1)
    errs = {}
    for name, prec, mf in (("fp64", "double", False), ("fp32", "float", False), ("fp32_mfma", "float", True)):
        x = x64.astype(np.complex128 if prec == "double" else np.complex64)
        a = fa.DeviceBuffer.from_array(x)
        b = fa.DeviceBuffer(x.nbytes)
        fl = fa.FFTW_ESTIMATE | (fa.FFTW_AMD_MFMA if mf else 0)
        p = fa.plan_many_dft(prec, [n], H, a.ptr, None, 1, n, b.ptr, None, 1, n, -1, fl)
        if mf:
            assert fa.lib().fftw_amd_plan_use_mfma(p.h) >= 1
        p.execute()
        y = b.download(x.dtype)
        errs[name] = oracle.rel_l2(y, ref)
    print(errs)
    assert errs["fp64"] <= 1e-13
    assert errs["fp32"] <= 2e-6
    assert errs["fp32_mfma"] <= 2e-6
