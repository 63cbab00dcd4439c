"""Rank >= 2 real transforms on the cross-row path: the row pass runs the
half-length r2c (or c2r) of the X rows of a tile together with the first
(or last) radix-X stage of the next dimension's DFT across them, and that
dimension's remaining (n0/X)-point DFT runs over rows X apart.  Each case is
compared with the oracle; the plan is checked to take that path."""
import numpy as np
import pytest

import fftw3_amd as fa
import oracle
from gpu_util import TOL, cdt, rdt

pytestmark = pytest.mark.gpu

# (prec, shape, howmany)
CASES = [("float", (64, 2048), 3), ("float", (16, 1024), 2), ("double", (24, 1024), 2),
         ("float", (2, 16, 2048), 2), ("float", (40, 1024), 1), ("double", (8 * 7, 1024), 1),
         ("float", (3, 16, 1024), 1)]


def uses_xr(p, fwd):
    s = p.sprint()
    return ("r2c-rows+radix-" in s) if fwd else ("+c2r-rows" in s)


@pytest.mark.parametrize("prec,shape,H", CASES)
def test_r2c_cross_row(eng, prec, shape, H):
    N = int(np.prod(shape))
    nh = shape[-1] // 2 + 1
    M = N // shape[-1] * nh
    x = oracle.uniform(N * H, 61).astype(rdt(prec))
    bi = fa.DeviceBuffer.from_array(x)
    bo = fa.DeviceBuffer(M * H * np.dtype(cdt(prec)).itemsize)
    p = fa.plan_many_dft_r2c(prec, list(shape), H, bi.ptr, None, 1, N, bo.ptr, None, 1, M)
    assert uses_xr(p, True), p.sprint()
    p.execute()
    y = bo.download(cdt(prec), M * H)
    ref = oracle.rfftn(x.astype(np.float64).reshape((H,) + shape), axes_batch=1).reshape(-1)
    err = oracle.rel_l2(y, ref)
    print(prec, shape, H, "r2c rel_l2 %.3e" % err)
    assert err <= TOL[prec], err


@pytest.mark.parametrize("prec,shape,H", CASES)
def test_c2r_cross_row(eng, prec, shape, H):
    N = int(np.prod(shape))
    nh = shape[-1] // 2 + 1
    M = N // shape[-1] * nh
    xr = oracle.uniform(N * H, 62).reshape((H,) + shape)
    X = oracle.rfftn(xr, axes_batch=1).astype(cdt(prec)).reshape(-1)
    bi = fa.DeviceBuffer.from_array(X)
    bo = fa.DeviceBuffer(N * H * np.dtype(rdt(prec)).itemsize)
    p = fa.plan_many_dft_c2r(prec, list(shape), H, bi.ptr, None, 1, M, bo.ptr, None, 1, N)
    assert uses_xr(p, False), p.sprint()
    p.execute()
    y = bo.download(rdt(prec), N * H)
    ref = oracle.irfftn(X.astype(np.complex128).reshape((H,) + shape[:-1] + (nh,)), shape[-1],
                        axes_batch=1).reshape(-1)
    err = oracle.rel_l2(y, ref)
    print(prec, shape, H, "c2r rel_l2 %.3e" % err)
    assert err <= TOL[prec], err


def test_padded_layout_and_measure(eng):
    """Rows padded in both arrays (inembed / onembed), FFTW_MEASURE (which
    times the cross-row and the plain two-pass forms); forward against the
    oracle, then the c2r round trip."""
    prec, n0, n1, H = "float", 128, 2048, 2
    nh = n1 // 2 + 1
    pin, pout = n1 + 8, nh + 3
    x = oracle.uniform(n0 * pin * H, 63).astype(np.float32)
    bi = fa.DeviceBuffer(x.nbytes)
    bo = fa.DeviceBuffer(n0 * pout * H * 8)
    br = fa.DeviceBuffer(x.nbytes)
    pf = fa.plan_many_dft_r2c(prec, [n0, n1], H, bi.ptr, [n0, pin], 1, n0 * pin, bo.ptr, [n0, pout], 1,
                              n0 * pout, fa.FFTW_MEASURE)
    pb = fa.plan_many_dft_c2r(prec, [n0, n1], H, bo.ptr, [n0, pout], 1, n0 * pout, br.ptr, [n0, pin], 1,
                              n0 * pin, fa.FFTW_MEASURE)
    print(pf.sprint(), pb.sprint())
    bi.upload(x)
    pf.execute()
    y = bo.download(np.complex64).reshape(H, n0, pout)[:, :, :nh]
    xs = x.reshape(H, n0, pin)[:, :, :n1].astype(np.float64)
    ref = oracle.rfftn(xs, axes_batch=1)
    assert oracle.rel_l2(y, ref) <= TOL[prec]
    pb.execute()
    z = br.download(np.float32).reshape(H, n0, pin)[:, :, :n1]
    assert oracle.rel_l2(z / (n0 * n1), xs) <= TOL[prec]
