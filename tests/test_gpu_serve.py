"""Serving loop (fftw_amd_server_*): continuous batching, admission by an
HBM budget, and results identical alone and in the mix.

"Alone" is the ordinary plan/execute API on the same input; "in the mix"
is the server with many requests of other sizes, kinds and precisions
admitted around it and in flight at the same time.
"""
import ctypes

import numpy as np
import pytest

import fftw3_amd as fa
import oracle
from gpu_util import TOL, cdt, rdt

pytestmark = pytest.mark.gpu

# (kind, prec, shape, howmany)
MIX = [("c2c", "float", (256,), 64), ("c2c", "float", (4096,), 8), ("c2c", "double", (1000,), 5),
       ("c2c", "float", (1 << 16,), 2), ("c2c", "float", (4093,), 3), ("r2c", "float", (64, 64), 4),
       ("c2c", "double", (32, 32, 32), 1), ("c2c", "float", (1 << 20,), 1), ("c2c", "double", (12,), 100),
       ("r2c", "double", (96, 128), 2), ("c2c", "float", (3 * 1024,), 4), ("c2c", "double", (65537,), 1)]


def make_input(kind, prec, shape, h, seed):
    N = int(np.prod(shape))
    if kind == "c2c":
        return oracle.uniform(2 * N * h, seed).astype(rdt(prec)).view(cdt(prec))
    return oracle.uniform(N * h, seed).astype(rdt(prec))


def out_len(kind, shape, h):
    if kind == "c2c":
        return int(np.prod(shape)) * h
    return int(np.prod(shape[:-1])) * (shape[-1] // 2 + 1) * h


def alone(kind, prec, shape, h, x):
    bi = fa.DeviceBuffer.from_array(x)
    n = out_len(kind, shape, h)
    bo = fa.DeviceBuffer(n * np.dtype(cdt(prec)).itemsize)
    N = int(np.prod(shape))
    if kind == "c2c":
        p = fa.plan_many_dft(prec, list(shape), h, bi.ptr, None, 1, N, bo.ptr, None, 1, N, -1)
    else:
        p = fa.plan_many_dft_r2c(prec, list(shape), h, bi.ptr, None, 1, N, bo.ptr, None, 1, n // h)
    p.execute()
    return bo.download(cdt(prec), n)


def oracle_ref(kind, shape, h, x):
    if kind == "c2c":
        return oracle.fftn(x.astype(np.complex128).reshape((h,) + shape), -1, axes_batch=1).reshape(-1)
    return oracle.rfftn(x.astype(np.float64).reshape((h,) + shape), axes_batch=1).reshape(-1)


def test_mix_equals_alone(eng):
    srv = fa.Server(2 << 30, max_rounds=4, per_round=16)
    ins, reqs, ids = [], [], []
    for rep in range(2):
        for i, (kind, prec, shape, h) in enumerate(MIX):
            x = make_input(kind, prec, shape, h, 100 + i)
            b = fa.DeviceBuffer.from_array(x)
            r = fa.Server.request(kind, prec, shape, h, b.ptr)
            ins.append((kind, prec, shape, h, x, b))
            ids.append(srv.submit(r))
    assert all(i >= 0 for i in ids)
    for rid, (kind, prec, shape, h, x, b) in zip(ids, ins):
        ptr, lat = srv.wait(rid)
        n = out_len(kind, shape, h)
        y = np.empty(n, cdt(prec))
        assert fa.lib().fftw_amd_memcpy_dtoh(y.ctypes.data, ptr, y.nbytes) == 0
        srv.release(rid)
        ya = alone(kind, prec, shape, h, x)
        assert np.array_equal(y, ya), (kind, prec, shape)
        assert oracle.rel_l2(y, oracle_ref(kind, shape, h, x)) <= TOL[prec], (kind, prec, shape)
        assert lat > 0
    st = srv.stats()
    assert st["completed"] == len(ids)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    assert st["hbm_in_use"] == 0
    srv.destroy()


def test_admission_by_budget(eng):
    # each request's output is 4096*8*8 = 256 KiB; the budget holds two
    n, h = 4096, 8
    out_bytes = n * h * 8
    srv = fa.Server(2 * out_bytes + 1024, max_rounds=4, per_round=8)
    xs, bufs, ids = [], [], []
    for i in range(6):
        x = make_input("c2c", "float", (n,), h, 200 + i)
        b = fa.DeviceBuffer.from_array(x)
        xs.append(x)
        bufs.append(b)
        ids.append(srv.submit(fa.Server.request("c2c", "float", (n,), h, b.ptr)))
    p0, _ = srv.wait(ids[0])
    p1, _ = srv.wait(ids[1])
    import time
    time.sleep(0.05)
    # the third cannot have been admitted: the budget is full until a release
    assert srv.poll(ids[2]) == 0
    assert srv.stats()["hbm_in_use"] >= 2 * out_bytes
    for k in range(6):
        ptr, _ = srv.wait(ids[k])
        y = np.empty(n * h, np.complex64)
        fa.lib().fftw_amd_memcpy_dtoh(y.ctypes.data, ptr, y.nbytes)
        ref = oracle_ref("c2c", (n,), h, xs[k])
        assert oracle.rel_l2(y, ref) <= TOL["float"]
        assert srv.release(ids[k]) == 0
    # larger than the whole budget: rejected at submission
    big = fa.DeviceBuffer(1 << 22)
    assert srv.submit(fa.Server.request("c2c", "float", (1 << 18,), 1, big.ptr)) == -1
    srv.destroy()


def test_invalid_requests(eng):
    srv = fa.Server(1 << 20)
    b = fa.DeviceBuffer(4096)
    assert srv.submit(fa.Server.request("c2c", "float", (0,), 1, b.ptr)) == -1
    assert srv.submit(fa.Server.request("c2c", "float", (8,), 0, b.ptr)) == -1
    assert srv.submit(fa.Server.request("c2c", "float", (8,), 1, 0)) == -1
    assert srv.poll(12345) == -1
    assert srv.release(777) == -1
    srv.destroy()
