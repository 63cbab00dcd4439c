"""Real-data DFT parity (r2c / c2r) against the oracle.

Layout semantics follow api/plan-many-dft-r2c.c:24-57 and
api/rdft2-pad.c:24-39 (n/2+1 complex outputs along the last dim; in-place
rows padded to 2(n/2+1) reals); c2r is the unnormalized inverse and, being
multi-dimensional, may destroy its input (api/plan-many-dft-c2r.c:41-42).
Inputs to c2r are the r2c outputs of real data, i.e. Hermitian, as the
reference's tests constrain them (libbench2/verify-lib.c:114-128).
"""
import numpy as np
import pytest

import fftw3_amd as fa
import oracle
from gpu_util import TOL, cdt, rdt, inputs_r

pytestmark = pytest.mark.gpu


def r2c_c2r(prec, shape, howmany, seed=2, inplace=False):
    shape = list(shape)
    nl = shape[-1]
    nh = nl // 2 + 1
    lead = int(np.prod(shape[:-1])) if len(shape) > 1 else 1
    creal = lead * nl
    ccplx = lead * nh
    x = inputs_r(creal * howmany, seed, prec).reshape(howmany, lead, nl)
    if inplace:
        # rows padded to 2*nh reals, complex output shares the buffer
        padded = np.zeros((howmany, lead, 2 * nh), rdt(prec))
        padded[:, :, :nl] = x
        buf = fa.DeviceBuffer.from_array(padded)
        bi = bo = buf
        idist, odist = lead * 2 * nh, ccplx
    else:
        bi = fa.DeviceBuffer.from_array(x)
        bo = fa.DeviceBuffer(ccplx * howmany * np.dtype(cdt(prec)).itemsize)
        idist, odist = creal, ccplx
    p = fa.plan_many_dft_r2c(prec, shape, howmany, bi.ptr, None, 1, idist, bo.ptr, None, 1, odist)
    p.execute()
    Y = bo.download(cdt(prec), ccplx * howmany).reshape(howmany, ccplx)
    ref = oracle.rfftn(x.astype(np.float64).reshape([howmany] + shape), axes_batch=1).reshape(howmany, -1)
    e_fwd = oracle.rel_l2(Y, ref)
    # c2r back (destroys its input when multi-dimensional)
    if inplace:
        q = fa.plan_many_dft_c2r(prec, shape, howmany, bo.ptr, None, 1, odist, bi.ptr, None, 1, idist)
        q.execute()
        z = bi.download(rdt(prec), howmany * lead * 2 * nh).reshape(howmany, lead, 2 * nh)[:, :, :nl]
    else:
        zb = fa.DeviceBuffer(creal * howmany * np.dtype(rdt(prec)).itemsize)
        q = fa.plan_many_dft_c2r(prec, shape, howmany, bo.ptr, None, 1, odist, zb.ptr, None, 1, idist)
        q.execute()
        z = zb.download(rdt(prec), howmany * creal).reshape(howmany, lead, nl)
    zref = oracle.irfftn(ref.reshape([howmany] + shape[:-1] + [nh]), nl, axes_batch=1)
    e_bwd = oracle.rel_l2(z.reshape(-1), zref.reshape(-1))
    e_rt = oracle.rel_l2(z.reshape(-1) / (lead * nl), x.reshape(-1).astype(np.float64))
    return e_fwd, e_bwd, e_rt


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("n", list(range(1, 65)) + [100, 128, 1000, 1001, 2048, 4096, 8192, 16384,
                                                   4093, 65536, 1 << 20, 130, 2018, 8186, 32768 * 3])
def test_r2c_c2r_1d(eng, prec, n):
    e_fwd, e_bwd, e_rt = r2c_c2r(prec, [n], 3, seed=n)
    assert e_fwd <= TOL[prec], e_fwd
    assert e_bwd <= TOL[prec], e_bwd
    assert e_rt <= 2 * TOL[prec], e_rt


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("shape", [(6, 8), (5, 9), (32, 32), (64, 48), (16, 8, 10), (2048, 2048)])
def test_r2c_c2r_multid(eng, prec, shape):
    H = 2 if np.prod(shape) > 100000 else 3
    e_fwd, e_bwd, e_rt = r2c_c2r(prec, shape, H, seed=2)
    assert e_fwd <= TOL[prec], e_fwd
    assert e_bwd <= TOL[prec], e_bwd
    assert e_rt <= 2 * TOL[prec], e_rt


@pytest.mark.parametrize("prec", ["double", "float"])
@pytest.mark.parametrize("shape", [(16,), (15,), (8, 12), (64, 64), (2018,), (4, 65536)])
def test_r2c_c2r_inplace(eng, prec, shape):
    e_fwd, e_bwd, e_rt = r2c_c2r(prec, shape, 2, seed=4, inplace=True)
    assert e_fwd <= TOL[prec], e_fwd
    assert e_bwd <= TOL[prec], e_bwd


def test_c2r_multid_preserve_input_is_null(eng):
    b = fa.DeviceBuffer(1 << 16)
    n = (fa.ctypes.c_int * 2)(8, 8)
    assert not fa.lib().fftw_plan_dft_c2r(2, n, b.ptr, b.ptr + 4096, fa.FFTW_PRESERVE_INPUT | fa.FFTW_ESTIMATE)


def test_r2r_planner_through_c_abi(eng):
    # plan_many_r2r is provided (every kind; numerics in test_gpu_r2r.py);
    # an unknown kind returns NULL like any unsolvable problem
    b = fa.DeviceBuffer(1024)
    n = (fa.ctypes.c_int * 1)(8)
    k = (fa.ctypes.c_int * 1)(0)
    p = fa.lib().fftw_plan_many_r2r(1, n, 1, b.ptr, None, 1, 8, b.ptr, None, 1, 8, k, fa.FFTW_ESTIMATE)
    assert p
    fa.lib().fftw_destroy_plan(p)
    k[0] = 42
    assert not fa.lib().fftw_plan_many_r2r(1, n, 1, b.ptr, None, 1, 8, b.ptr, None, 1, 8, k, fa.FFTW_ESTIMATE)


@pytest.mark.parametrize("prec", ["double", "float"])
def test_c2r_1d_ignores_imag_dc_nyquist(eng, prec):
    n = 64
    X = oracle.uniform_complex(n // 2 + 1, 21)
    Xd = X.astype(cdt(prec))
    b = fa.DeviceBuffer.from_array(Xd)
    o = fa.DeviceBuffer(n * np.dtype(rdt(prec)).itemsize)
    p = fa.plan_many_dft_c2r(prec, [n], 1, b.ptr, None, 1, n // 2 + 1, o.ptr, None, 1, n)
    p.execute()
    y = o.download(rdt(prec))
    ref = oracle.irfftn(Xd.astype(np.complex128), n)
    assert oracle.rel_l2(y, ref) <= TOL[prec]
