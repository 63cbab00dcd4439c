"""Committed oracle fixtures (tests/golden/golden.npz).

CPU: the oracle still reproduces its frozen outputs (guards the oracle itself).
GPU: the engine matches the frozen outputs through the C ABI, with no live
oracle involved.
"""
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def golden():
    z = np.load(os.path.join(HERE, "golden", "golden.npz"))
    meta = json.loads(bytes(z["meta"]).decode())
    return z, meta


def test_oracle_reproduces_fixtures(golden):
    z, meta = golden
    for n, seed in meta["c2c"]:
        x = oracle.uniform_complex(n, seed)
        assert oracle.rel_l2(oracle.fftn(x, -1), z["c2c_fwd_%d" % n]) < 1e-16
        assert oracle.rel_l2(oracle.fftn(x, +1), z["c2c_bwd_%d" % n]) < 1e-16
    for n, seed in meta["r2c"]:
        x = oracle.uniform(n, seed)
        assert oracle.rel_l2(oracle.rfftn(x), z["r2c_%d" % n]) < 1e-16
    for n, seed in meta["r2r"]:
        x = oracle.uniform(n, seed)
        for kind in range(11):
            assert oracle.rel_l2(oracle.r2rn(x, [kind]), z["r2r_%d_%d" % (kind, n)]) < 1e-16
    for shape, kinds, seed in meta["r2r_multi"]:
        x = oracle.uniform(int(np.prod(shape)), seed).reshape(shape)
        key = "r2r_multi_%s" % "x".join(map(str, shape))
        assert oracle.rel_l2(oracle.r2rn(x, kinds), z[key]) < 1e-16


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["double", "float"])
def test_engine_matches_fixtures(eng, golden, prec):
    from gpu_util import TOL, cdt, rdt
    z, meta = golden
    fa = eng
    for n, seed in meta["c2c"]:
        x = oracle.uniform_complex(n, seed).astype(cdt(prec))
        for sign, key in ((-1, "c2c_fwd_%d"), (1, "c2c_bwd_%d")):
            a = fa.DeviceBuffer.from_array(x)
            b = fa.DeviceBuffer(x.nbytes)
            p = fa.plan_many_dft(prec, [n], 1, a.ptr, None, 1, n, b.ptr, None, 1, n, sign)
            p.execute()
            y = b.download(cdt(prec))
            ref = z[key % n]
            # the fixture is computed from the FP64 input; FP32 inputs are
            # rounded, which the FP32 bar absorbs
            assert oracle.rel_l2(y, ref) <= (TOL[prec] if prec == "double" else 5e-6), (n, sign)
    for n, seed in meta["r2c"]:
        x = oracle.uniform(n, seed).astype(rdt(prec))
        a = fa.DeviceBuffer.from_array(x)
        b = fa.DeviceBuffer((n // 2 + 1) * np.dtype(cdt(prec)).itemsize)
        p = fa.plan_many_dft_r2c(prec, [n], 1, a.ptr, None, 1, n, b.ptr, None, 1, n // 2 + 1)
        p.execute()
        assert oracle.rel_l2(b.download(cdt(prec)), z["r2c_%d" % n]) <= (TOL[prec] if prec == "double" else 5e-6)
    for shape, seed in meta["multi"]:
        N = int(np.prod(shape))
        x = oracle.uniform_complex(N, seed).astype(cdt(prec))
        a = fa.DeviceBuffer.from_array(x)
        b = fa.DeviceBuffer(x.nbytes)
        p = fa.plan_many_dft(prec, list(shape), 1, a.ptr, None, 1, N, b.ptr, None, 1, N, -1)
        p.execute()
        ref = z["multi_%s" % "x".join(map(str, shape))].reshape(-1)
        assert oracle.rel_l2(b.download(cdt(prec)), ref) <= (TOL[prec] if prec == "double" else 5e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["double", "float"])
def test_engine_matches_r2r_fixtures(eng, golden, prec):
    from gpu_util import TOL, rdt
    z, meta = golden
    fa = eng
    bar = TOL[prec] if prec == "double" else 5e-6
    for n, seed in meta["r2r"]:
        x = oracle.uniform(n, seed).astype(rdt(prec))
        a = fa.DeviceBuffer.from_array(x)
        b = fa.DeviceBuffer(x.nbytes)
        for kind in range(11):
            p = fa.plan_many_r2r(prec, [n], 1, a.ptr, None, 1, n, b.ptr, None, 1, n, [kind])
            p.execute()
            assert oracle.rel_l2(b.download(rdt(prec)), z["r2r_%d_%d" % (kind, n)]) <= bar, (kind, n)
            p.destroy()
    for shape, kinds, seed in meta["r2r_multi"]:
        N = int(np.prod(shape))
        x = oracle.uniform(N, seed).astype(rdt(prec))
        a = fa.DeviceBuffer.from_array(x)
        b = fa.DeviceBuffer(x.nbytes)
        p = fa.plan_many_r2r(prec, list(shape), 1, a.ptr, None, 1, N, b.ptr, None, 1, N, kinds)
        p.execute()
        key = "r2r_multi_%s" % "x".join(map(str, shape))
        assert oracle.rel_l2(b.download(rdt(prec)), z[key].reshape(-1)) <= bar, shape
        p.destroy()
