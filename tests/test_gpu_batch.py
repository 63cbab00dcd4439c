"""Batched uneven sizes in persistent launches (fftw_amd_plan_batch_dft_1d).

Every group's result must be identical, bit for bit, alone (its own
fftw plan, ordinary kernels) and in the mix (persistent work-table launches
replayed through a hipGraph), and within the parity bar of the oracle.
Graph replays are checked one after another with eager work in between.
"""
import numpy as np
import pytest

import fftw3_amd as fa
import oracle
from gpu_util import TOL, cdt, inputs_c

pytestmark = pytest.mark.gpu

MIX = [(1000, 3), (1001, 2), (2187, 2), (4093, 5), (4096, 4), (65536, 2), (1 << 20, 1),
       (97, 6), (256, 7), (12, 9)]


def setup_groups(prec, mix, seed0=30):
    groups, xs, bufs = [], [], []
    for i, (n, h) in enumerate(mix):
        x = inputs_c(n * h, seed0 + i, prec)
        bi = fa.DeviceBuffer.from_array(x)
        bo = fa.DeviceBuffer(x.nbytes)
        groups.append((n, h, bi.ptr, bo.ptr))
        xs.append(x)
        bufs.append((bi, bo))
    return groups, xs, bufs


def alone(prec, n, h, x, sign):
    bi = fa.DeviceBuffer.from_array(x)
    bo = fa.DeviceBuffer(x.nbytes)
    p = fa.plan_many_dft(prec, [n], h, bi.ptr, None, 1, n, bo.ptr, None, 1, n, sign)
    p.execute()
    return bo.download(cdt(prec))


@pytest.mark.parametrize("prec", ["float", "double"])
@pytest.mark.parametrize("sign", [-1, 1])
def test_mix_equals_alone_bitwise(eng, prec, sign):
    groups, xs, bufs = setup_groups(prec, MIX)
    b = fa.Batch(prec, groups, sign)
    print(b.info(), flush=True)
    b.execute()
    for (n, h, _, _), x, (bi, bo) in zip(groups, xs, bufs):
        y = bo.download(cdt(prec))
        ya = alone(prec, n, h, x, sign)
        assert np.array_equal(y.view(np.uint8), ya.view(np.uint8)), n
        ref = oracle.fftn(x.astype(np.complex128).reshape(h, n), sign, axes_batch=1).reshape(-1)
        assert oracle.rel_l2(y, ref) <= TOL[prec], n


def test_graph_replays_and_direct_launches_agree(eng):
    prec = "float"
    groups, xs, bufs = setup_groups(prec, MIX[:7], seed0=50)
    b = fa.Batch(prec, groups, -1)
    outs = []
    for rep in range(3):
        # eager work between replays: refresh inputs, clear outputs
        for (n, h, _, _), x, (bi, bo) in zip(groups, xs, bufs):
            bi.upload(x)
            bo.upload(np.zeros(n * h, cdt(prec)))
        fa.lib().fftw_amd_synchronize()
        b.execute(graph=True)
        outs.append([bo.download(cdt(prec)) for (_, bo) in bufs])
    for (_, bo) in bufs:
        bo.upload(np.zeros(bo.nbytes // 8, cdt(prec)))
    b.execute(graph=False)
    direct = [bo.download(cdt(prec)) for (_, bo) in bufs]
    for rep in range(3):
        for a, d in zip(outs[rep], direct):
            assert np.array_equal(a, d)


def test_batch_in_place_and_measure(eng):
    prec = "float"
    mix = [(4096, 8), (1024, 16), (65536, 2)]
    groups, xs = [], []
    keep = []
    for i, (n, h) in enumerate(mix):
        x = inputs_c(n * h, 70 + i, prec)
        buf = fa.DeviceBuffer.from_array(x)
        keep.append(buf)
        groups.append((n, h, buf.ptr, buf.ptr))
        xs.append(x)
    b = fa.Batch(prec, groups, -1, fa.FFTW_MEASURE)
    for (n, h, p, _), x, buf in zip(groups, xs, keep):
        buf.upload(x)  # FFTW_MEASURE may overwrite the arrays
    b.execute()
    for (n, h, p, _), x, buf in zip(groups, xs, keep):
        ref = oracle.fftn(x.astype(np.complex128).reshape(h, n), -1, axes_batch=1).reshape(-1)
        assert oracle.rel_l2(buf.download(cdt(prec)), ref) <= TOL[prec], n


def test_batch_rejects_bad_groups(eng):
    b = fa.DeviceBuffer(4096)
    with pytest.raises(RuntimeError):
        fa.Batch("float", [(0, 1, b.ptr, b.ptr)])
    with pytest.raises(RuntimeError):
        fa.Batch("float", [(8, -1, b.ptr, b.ptr)])
