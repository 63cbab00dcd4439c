"""Cache-blocked multi-dimensional plans (FFTW_AMD_BLOCKED): batched
problems run every pass per chunk of transforms, unbatched rank-3 problems
run the inner-dimension passes per slab of planes.  The kernels are the
same as in the whole-problem plan, so the blocked result must equal the
unblocked one bit for bit; the oracle checks the transform itself on a
sample (rank-split semantics of dft/rank-geq2.c:42-191)."""
import numpy as np
import pytest

import fftw3_amd as fa
import oracle
from gpu_util import TOL, cdt, inputs_c, inputs_r, rdt

pytestmark = pytest.mark.gpu


def blocked(p):
    return "blocked" in p.sprint()


@pytest.mark.parametrize("prec,shape", [("double", (128, 256, 256)), ("float", (256, 256, 256))])
def test_blocked_3d_slabs(prec, shape):
    N = int(np.prod(shape))
    x = inputs_c(N, 31, prec)
    a = fa.DeviceBuffer.from_array(x)
    ref_buf = fa.DeviceBuffer.from_array(x)
    for sign in (-1, 1):
        a.upload(x)
        ref_buf.upload(x)
        pb = fa.plan_many_dft(prec, list(shape), 1, a.ptr, None, 1, N, a.ptr, None, 1, N, sign,
                              fa.FFTW_ESTIMATE | fa.FFTW_AMD_BLOCKED)
        pu = fa.plan_many_dft(prec, list(shape), 1, ref_buf.ptr, None, 1, N, ref_buf.ptr, None, 1, N,
                              sign, fa.FFTW_ESTIMATE)
        assert blocked(pb) and not blocked(pu)
        pb.execute()
        pu.execute()
        yb = a.download(cdt(prec))
        yu = ref_buf.download(cdt(prec))
        assert np.array_equal(yb, yu)
        # oracle on sampled output bins: Y[k0,k1,k2] = sum x e^{-+2 pi i k.j / n}
        xs = x.astype(np.complex128).reshape(shape)
        rng = np.random.default_rng(3)
        ks = [tuple(int(rng.integers(0, n)) for n in shape) for _ in range(4)]
        for k in ks:
            ph = np.ones(shape, np.complex128)
            for d, n in enumerate(shape):
                w = np.exp(sign * 2j * np.pi * k[d] * np.arange(n) / n)
                sh = [1] * len(shape)
                sh[d] = n
                ph = ph * w.reshape(sh)
            want = np.sum(xs * ph)
            got = yb.reshape(shape)[k]
            assert abs(got - want) / np.sqrt(N) < 20 * TOL[prec] * np.sqrt(np.log2(N))
        pb.destroy()
        pu.destroy()


@pytest.mark.parametrize("prec", ["double", "float"])
def test_blocked_batched_c2c_2d(prec):
    shape, H = (512, 512), 48 if prec == "double" else 96
    N = 512 * 512
    x = inputs_c(N * H, 32, prec)
    a = fa.DeviceBuffer.from_array(x)
    o1, o2 = fa.DeviceBuffer(x.nbytes), fa.DeviceBuffer(x.nbytes)
    pb = fa.plan_many_dft(prec, list(shape), H, a.ptr, None, 1, N, o1.ptr, None, 1, N, -1,
                          fa.FFTW_ESTIMATE | fa.FFTW_AMD_BLOCKED)
    pu = fa.plan_many_dft(prec, list(shape), H, a.ptr, None, 1, N, o2.ptr, None, 1, N, -1,
                          fa.FFTW_ESTIMATE)
    assert blocked(pb)
    pb.execute()
    pu.execute()
    y1, y2 = o1.download(cdt(prec)), o2.download(cdt(prec))
    assert np.array_equal(y1, y2)
    for h in (0, H - 1):
        ref = oracle.fftn(x[h * N:(h + 1) * N].astype(np.complex128).reshape(shape), -1).reshape(-1)
        assert oracle.rel_l2(y1[h * N:(h + 1) * N], ref) < TOL[prec]
    pb.destroy()
    pu.destroy()


@pytest.mark.parametrize("prec", ["double", "float"])
def test_blocked_r2c_c2r_2d(prec):
    n0, n1, H = 1024, 1024, 16 if prec == "double" else 32
    N, nh = n0 * n1, n1 // 2 + 1
    M = n0 * nh
    x = inputs_r(N * H, 33, prec)
    a = fa.DeviceBuffer.from_array(x)
    c1, c2 = fa.DeviceBuffer(M * H * 2 * x.itemsize), fa.DeviceBuffer(M * H * 2 * x.itemsize)
    pb = fa.plan_many_dft_r2c(prec, [n0, n1], H, a.ptr, None, 1, N, c1.ptr, None, 1, M,
                              fa.FFTW_ESTIMATE | fa.FFTW_AMD_BLOCKED)
    pu = fa.plan_many_dft_r2c(prec, [n0, n1], H, a.ptr, None, 1, N, c2.ptr, None, 1, M, fa.FFTW_ESTIMATE)
    assert blocked(pb)
    pb.execute()
    pu.execute()
    X1, X2 = c1.download(cdt(prec)), c2.download(cdt(prec))
    assert np.array_equal(X1, X2)
    for h in (0, H - 1):
        ref = oracle.rfftn(x[h * N:(h + 1) * N].astype(np.float64).reshape(n0, n1)).reshape(-1)
        assert oracle.rel_l2(X1[h * M:(h + 1) * M], ref) < TOL[prec]
    # c2r (destroys its input): blocked and unblocked on equal copies
    r1, r2 = fa.DeviceBuffer(x.nbytes), fa.DeviceBuffer(x.nbytes)
    qb = fa.plan_many_dft_c2r(prec, [n0, n1], H, c1.ptr, None, 1, M, r1.ptr, None, 1, N,
                              fa.FFTW_ESTIMATE | fa.FFTW_AMD_BLOCKED)
    qu = fa.plan_many_dft_c2r(prec, [n0, n1], H, c2.ptr, None, 1, M, r2.ptr, None, 1, N, fa.FFTW_ESTIMATE)
    assert blocked(qb)
    qb.execute()
    qu.execute()
    y1, y2 = r1.download(rdt(prec)), r2.download(rdt(prec))
    assert np.array_equal(y1, y2)
    assert oracle.rel_l2(y1[:N].astype(np.float64) / N, x[:N].astype(np.float64)) < TOL[prec]
    for p in (pb, pu, qb, qu):
        p.destroy()
