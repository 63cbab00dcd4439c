"""The allocation-free execute (fftw_amd_execute_noalloc) against the
ordinary new-array execute of the same plan.

Each problem's plan uses scratch (a pass whose source or destination is a
scratch slot).  The allocation-free call must give bit-identical output on
two calls in a row, the second with a smaller problem, so that whatever the
first call left in the shared static scratch cannot leak into the second;
both are also within the oracle bar.  Plans whose scratch exceeds the static
size are rejected with an error, not run.
"""
import ctypes

import numpy as np
import pytest

import fftw3_amd as fa
import oracle
from gpu_util import TOL, cdt, rdt

pytestmark = pytest.mark.gpu

STATIC_BYTES = 128 << 20


def noalloc(p, i, o):
    L = fa.lib()
    L.fftw_amd_execute_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.fftw_amd_execute_noalloc.restype = ctypes.c_int
    return L.fftw_amd_execute_noalloc(p.h, i, o)


def uses_scratch(p):
    return any(" src=s" in s or " dst=s" in s for s in p.pass_info())


def c2c_case(prec, n, h, seed):
    x = oracle.uniform(2 * n * h, seed).astype(rdt(prec)).view(cdt(prec))
    bi = fa.DeviceBuffer.from_array(x)
    bo = fa.DeviceBuffer(x.nbytes)
    p = fa.plan_many_dft(prec, [n], h, bi.ptr, None, 1, n, bo.ptr, None, 1, n, -1)
    ref = oracle.fftn(x.astype(np.complex128).reshape(h, n), -1, axes_batch=1).reshape(-1)
    return p, bi, bo, x, ref, cdt(prec), x.size


def r2c_case(prec, n, h, seed):
    x = oracle.uniform(n * h, seed).astype(rdt(prec))
    bi = fa.DeviceBuffer.from_array(x)
    m = n // 2 + 1
    bo = fa.DeviceBuffer(m * h * np.dtype(cdt(prec)).itemsize)
    p = fa.plan_many_dft_r2c(prec, [n], h, bi.ptr, None, 1, n, bo.ptr, None, 1, m)
    ref = oracle.rfftn(x.astype(np.float64).reshape(h, n), axes_batch=1).reshape(-1)
    return p, bi, bo, x, ref, cdt(prec), m * h


# (maker, prec, first problem, second smaller problem)
CASES = [
    (c2c_case, "double", (1 << 20, 4), (1 << 18, 2)),
    (c2c_case, "float", (1 << 22, 2), (1 << 20, 1)),
    (c2c_case, "double", (65537, 2), (10007, 3)),
    (r2c_case, "float", (1 << 21, 2), (1 << 19, 1)),
]


@pytest.mark.parametrize("maker,prec,first,second", CASES)
def test_noalloc_matches_execute(eng, maker, prec, first, second):
    for k, (n, h) in enumerate((first, second)):
        p, bi, bo, x, ref, dt, cnt = maker(prec, n, h, 700 + 10 * k + n % 7)
        print("noalloc", prec, n, h, p.pass_info(), flush=True)
        assert uses_scratch(p), p.pass_info()
        # the ordinary execute of the same plan on the same arrays
        p.execute()
        want = bo.download(dt, cnt)
        bo.upload(np.zeros(cnt, dt))
        bi.upload(x)
        assert noalloc(p, bi.ptr, bo.ptr) == 0, fa.last_error()
        got = bo.download(dt, cnt)
        assert np.array_equal(got, want), (prec, n, h)
        assert oracle.rel_l2(got, ref) <= TOL[prec], (prec, n, h)
        p.destroy()


def test_noalloc_rejects_oversized_scratch(eng):
    # a four-step FP64 c2c of 2^22 x 8 needs 512 MiB of scratch
    n, h = 1 << 22, 8
    bi = fa.DeviceBuffer(n * h * 16)
    bo = fa.DeviceBuffer(n * h * 16)
    p = fa.plan_many_dft("double", [n], h, bi.ptr, None, 1, n, bo.ptr, None, 1, n, -1)
    if not uses_scratch(p):
        pytest.skip("plan has no scratch: " + str(p.pass_info()))
    assert noalloc(p, bi.ptr, bo.ptr) != 0
    assert "static scratch" in fa.last_error()
    p.destroy()
