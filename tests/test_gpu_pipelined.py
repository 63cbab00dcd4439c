"""Pipelined row kernels (stockham_pipe): the grid is the resident
workgroups and each runs several tiles, loading the next tile's first-stage
inputs during the current tile's stages.  The catalog sweep runs them on a
few lines only (one tile per workgroup); these batches make every workgroup
loop, with ragged last rounds, both signs, in place and out of place, and
check every line against the oracle (the FP64 bar, 1e-13), plus a
bit-for-bit comparison with the one-tile kernel of the same geometry."""
import numpy as np
import pytest

import fftw3_amd as fa
import oracle
from gpu_util import TOL, c2c_many, cdt, inputs_c

pytestmark = pytest.mark.gpu


def plan_grid(p):
    s = p.pass_info()[0]
    assert " pipelined " in s, s
    return int(s.split(" grid=")[1].split()[0])


@pytest.mark.parametrize("n,H", [(4096, 1), (4096, 700), (4096, 1061), (2048, 2500)])
@pytest.mark.parametrize("sign", [-1, 1])
@pytest.mark.parametrize("inplace", [False, True])
def test_pipelined_rows_fp64(eng, n, H, sign, inplace):
    y, ref, _ = c2c_many("double", [n], H, sign, seed=40 + H, inplace=inplace)
    assert oracle.rel_l2(y, ref) < TOL["double"]
    # every line, not only the norm over the batch
    yl, rl = y.reshape(H, n), ref.reshape(H, n)
    err = np.linalg.norm(yl - rl, axis=1) / np.linalg.norm(rl, axis=1)
    assert err.max() < TOL["double"], int(err.argmax())


def test_pipelined_rounds_and_twin(eng):
    """A batch of more than two rounds of resident workgroups: the pipelined
    default and the one-tile kernel of the same geometry give identical
    output."""
    n = 4096
    cat = fa.catalog()
    pipe = [e for e in cat if e["prec"] == "double" and e["L"] == n and e["mode"] == "row"
            and e["io"] == "plain" and e["loop"]]
    assert len(pipe) == 1 and pipe[0]["variant"] == 0
    twin = [e for e in cat if e["prec"] == "double" and e["L"] == n and e["mode"] == "row"
            and e["io"] == "plain" and not e["loop"] and e["tpl"] == pipe[0]["tpl"]
            and e["nl"] == pipe[0]["nl"] and e["rad"] == pipe[0]["rad"]]
    assert len(twin) == 1
    x = inputs_c(n * 4096, 77, "double")
    a = fa.DeviceBuffer.from_array(x)
    b1, b2 = fa.DeviceBuffer(x.nbytes), fa.DeviceBuffer(x.nbytes)
    p1 = fa.plan_many_dft("double", [n], 4096, a.ptr, None, 1, n, b1.ptr, None, 1, n, -1,
                          fa.FFTW_ESTIMATE)
    grid = plan_grid(p1)
    assert 4096 > 2 * grid, grid
    fa.force_kernel(twin[0]["index"])
    try:
        p2 = fa.plan_many_dft("double", [n], 4096, a.ptr, None, 1, n, b2.ptr, None, 1, n, -1,
                              fa.FFTW_ESTIMATE)
    finally:
        fa.force_kernel(-1)
    assert " pipelined " not in p2.pass_info()[0]
    p1.execute()
    p2.execute()
    y1, y2 = b1.download(cdt("double")), b2.download(cdt("double"))
    assert np.array_equal(y1, y2)
    lines = [0, 1, 2047, 4095]
    ref = oracle.fftn(x.reshape(4096, n)[lines].astype(np.complex128), -1, axes_batch=1)
    assert oracle.rel_l2(y1.reshape(4096, n)[lines], ref) < TOL["double"]
    p1.destroy()
    p2.destroy()
