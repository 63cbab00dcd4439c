"""Split-complex guru interface on the GPU (api/plan-guru-split-dft.c,
api/plan-guru-split-dft-r2c.c, api/plan-guru-split-dft-c2r.c,
api/execute-split-dft*.c): real and imaginary parts in separate arrays,
strides counting reals, forward on (ri, ii) and backward by swapping the
pointers of each pair.  Checked against the long-double oracle on the same
seeded inputs, with the parity bars of gpu_util.TOL."""
import numpy as np
import pytest

import fftw3_amd as fa
import oracle
from gpu_util import TOL, cdt, inputs_c, inputs_r, rdt

pytestmark = pytest.mark.gpu

PRECS = ["double", "float"]


def row_major(shape, elem=1):
    """guru dims of a C-contiguous array, strides in elements x elem."""
    dims, s = [], elem
    for n in reversed(shape):
        dims.append((n, s, s))
        s *= n
    return list(reversed(dims)), s


def split_upload(x, prec):
    re = fa.DeviceBuffer.from_array(np.ascontiguousarray(x.real.astype(rdt(prec))))
    im = fa.DeviceBuffer.from_array(np.ascontiguousarray(x.imag.astype(rdt(prec))))
    return re, im


def split_download(re, im, prec, count):
    return re.download(rdt(prec), count).astype(np.float64) + 1j * im.download(rdt(prec), count)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("shape", [(64,), (1000,), (32, 48), (8, 12, 10), (4093,)])
def test_split_c2c_separate_arrays(prec, shape):
    H = 3
    N = int(np.prod(shape))
    x = inputs_c(N * H, 21, prec)
    dims, tot = row_major(shape)
    hd = [(H, tot, tot)]
    ri, ii = split_upload(x, prec)
    ro, io = fa.DeviceBuffer(x.real.nbytes), fa.DeviceBuffer(x.real.nbytes)
    xr = x.astype(np.complex128).reshape((H,) + shape)
    for sign in (-1, 1):
        if sign == -1:
            p = fa.plan_guru_split_dft(prec, dims, hd, ri.ptr, ii.ptr, ro.ptr, io.ptr)
        else:
            p = fa.plan_guru_split_dft(prec, dims, hd, ii.ptr, ri.ptr, io.ptr, ro.ptr)
        assert any(s.startswith("elementwise") for s in p.pass_info())  # staged path
        p.execute()
        y = split_download(ro, io, prec, N * H)
        ref = oracle.fftn(xr, sign, axes_batch=1).reshape(-1)
        assert oracle.rel_l2(y, ref) < TOL[prec], (sign, oracle.rel_l2(y, ref))
        # input arrays untouched (out of place)
        assert np.array_equal(split_download(ri, ii, prec, N * H), x.astype(np.complex128))
        p.destroy()


@pytest.mark.parametrize("prec", PRECS)
def test_split_c2c_strided_and_inplace(prec):
    # 1-D n=256, input stride 3 reals, output stride 2 reals, 4 transforms
    n, H, si, so, di, do = 256, 4, 3, 2, 800, 600
    span_i, span_o = di * H, do * H
    xre = inputs_r(span_i, 5, prec)
    xim = inputs_r(span_i, 6, prec)
    ri, ii = fa.DeviceBuffer.from_array(xre), fa.DeviceBuffer.from_array(xim)
    ro = fa.DeviceBuffer.from_array(np.zeros(span_o, rdt(prec)))
    io = fa.DeviceBuffer.from_array(np.zeros(span_o, rdt(prec)))
    p = fa.plan_guru_split_dft(prec, [(n, si, so)], [(H, di, do)], ri.ptr, ii.ptr, ro.ptr, io.ptr)
    p.execute()
    yre, yim = ro.download(rdt(prec)), io.download(rdt(prec))
    for h in range(H):
        xin = (xre[h * di + si * np.arange(n)].astype(np.float64)
               + 1j * xim[h * di + si * np.arange(n)].astype(np.float64))
        ref = oracle.fftn(xin, -1)
        got = yre[h * do + so * np.arange(n)] + 1j * yim[h * do + so * np.arange(n)].astype(np.float64)
        assert oracle.rel_l2(got, ref) < TOL[prec]
    # positions between the strided outputs are never written
    mask = np.zeros(span_o, bool)
    for h in range(H):
        mask[h * do + so * np.arange(n)] = True
    assert not np.any(yre[~mask]) and not np.any(yim[~mask])
    p.destroy()
    # in place on separate arrays: ri == ro, ii == io
    shape = (16, 40)
    dims, tot = row_major(shape)
    x = inputs_c(tot * 2, 7, prec)
    ri, ii = split_upload(x, prec)
    p = fa.plan_guru_split_dft(prec, dims, [(2, tot, tot)], ri.ptr, ii.ptr, ri.ptr, ii.ptr)
    p.execute()
    y = split_download(ri, ii, prec, tot * 2)
    ref = oracle.fftn(x.astype(np.complex128).reshape((2,) + shape), -1, axes_batch=1).reshape(-1)
    assert oracle.rel_l2(y, ref) < TOL[prec]
    p.destroy()


@pytest.mark.parametrize("prec", PRECS)
def test_split_paired_runs_interleaved(prec):
    """re/im views of one interleaved array: the plan has no staging passes;
    the swapped pairing is the backward transform."""
    shape, H = (512,), 8
    N = 512
    x = inputs_c(N * H, 8, prec)
    b = fa.DeviceBuffer.from_array(x)
    o = fa.DeviceBuffer(x.nbytes)
    rs = np.dtype(rdt(prec)).itemsize
    dims, tot = row_major(shape, 2)
    hd = [(H, 2 * N, 2 * N)]
    xr = x.astype(np.complex128).reshape(H, N)
    for sign in (-1, 1):
        if sign == -1:
            p = fa.plan_guru_split_dft(prec, dims, hd, b.ptr, b.ptr + rs, o.ptr, o.ptr + rs)
        else:
            p = fa.plan_guru_split_dft(prec, dims, hd, b.ptr + rs, b.ptr, o.ptr + rs, o.ptr)
        assert not any(s.startswith("elementwise") for s in p.pass_info())
        p.execute()
        y = o.download(cdt(prec))
        ref = oracle.fftn(xr, sign, axes_batch=1).reshape(-1)
        assert oracle.rel_l2(y, ref) < TOL[prec]
        p.destroy()


@pytest.mark.parametrize("prec", PRECS)
def test_execute_split_new_arrays(prec):
    shape, H = (24, 36), 2
    dims, tot = row_major(shape)
    x1 = inputs_c(tot * H, 9, prec)
    x2 = inputs_c(tot * H, 10, prec)
    a_r, a_i = split_upload(x1, prec)
    b_r, b_i = split_upload(x2, prec)
    oa_r, oa_i = fa.DeviceBuffer(a_r.nbytes), fa.DeviceBuffer(a_r.nbytes)
    ob_r, ob_i = fa.DeviceBuffer(a_r.nbytes), fa.DeviceBuffer(a_r.nbytes)
    p = fa.plan_guru_split_dft(prec, dims, [(H, tot, tot)], a_r.ptr, a_i.ptr, oa_r.ptr, oa_i.ptr)
    p.execute_split(b_r.ptr, b_i.ptr, ob_r.ptr, ob_i.ptr)
    y = split_download(ob_r, ob_i, prec, tot * H)
    ref = oracle.fftn(x2.astype(np.complex128).reshape((H,) + shape), -1, axes_batch=1).reshape(-1)
    assert oracle.rel_l2(y, ref) < TOL[prec]
    p.destroy()


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("shape", [(256,), (30, 64), (12, 18)])
def test_split_r2c_c2r(prec, shape):
    H = 2
    N = int(np.prod(shape))
    nh = shape[-1] // 2 + 1
    cshape = shape[:-1] + (nh,)
    M = int(np.prod(cshape))
    x = inputs_r(N * H, 11, prec)
    bx = fa.DeviceBuffer.from_array(x)
    ro, io = fa.DeviceBuffer(M * H * x.itemsize), fa.DeviceBuffer(M * H * x.itemsize)
    rdims, rt = row_major(shape)
    cdims, ct = row_major(cshape)
    dims = [(n, a[1], b[1]) for (n, a, b) in zip(shape, rdims, cdims)]
    p = fa.plan_guru_split_dft_r2c(prec, dims, [(H, rt, ct)], bx.ptr, ro.ptr, io.ptr)
    p.execute()
    X = split_download(ro, io, prec, M * H)
    ref = oracle.rfftn(x.astype(np.float64).reshape((H,) + shape), axes_batch=1).reshape(-1)
    assert oracle.rel_l2(X, ref) < TOL[prec]
    p.destroy()
    # c2r from the split spectrum; the staged plan leaves its input intact
    out = fa.DeviceBuffer(x.nbytes)
    dims_b = [(n, b[1], a[1]) for (n, a, b) in zip(shape, rdims, cdims)]
    q = fa.plan_guru_split_dft_c2r(prec, dims_b, [(H, ct, rt)], ro.ptr, io.ptr, out.ptr)
    q.execute()
    y = out.download(rdt(prec)).astype(np.float64)
    refb = oracle.irfftn(ref.reshape((H,) + cshape), shape[-1], axes_batch=1).reshape(-1)
    assert oracle.rel_l2(y, refb) < TOL[prec]
    assert np.array_equal(split_download(ro, io, prec, M * H), X)
    q.destroy()


@pytest.mark.parametrize("prec", PRECS)
def test_split_r2c_paired(prec):
    n, H = 1024, 4
    nh = n // 2 + 1
    x = inputs_r(n * H, 12, prec)
    bx = fa.DeviceBuffer.from_array(x)
    bo = fa.DeviceBuffer(nh * H * 2 * x.itemsize)
    rs = x.itemsize
    p = fa.plan_guru_split_dft_r2c(prec, [(n, 1, 2)], [(H, n, 2 * nh)], bx.ptr, bo.ptr, bo.ptr + rs)
    assert not any(s.startswith("elementwise") for s in p.pass_info())
    p.execute_split_r2c(bx.ptr, bo.ptr, bo.ptr + rs)
    X = bo.download(cdt(prec))
    ref = oracle.rfftn(x.astype(np.float64).reshape(H, n), axes_batch=1).reshape(-1)
    assert oracle.rel_l2(X, ref) < TOL[prec]
    p.destroy()


@pytest.mark.parametrize("prec", PRECS)
def test_split_rank0_and_guru64(prec):
    # rank 0: a strided copy of split arrays
    H1, H2 = 5, 7
    x = inputs_c(H1 * H2, 13, prec)
    ri, ii = split_upload(x, prec)
    ro, io = fa.DeviceBuffer(ri.nbytes), fa.DeviceBuffer(ri.nbytes)
    p = fa.plan_guru_split_dft(prec, [], [(H1, H2, 1), (H2, 1, H1)], ri.ptr, ii.ptr, ro.ptr, io.ptr)
    p.execute()
    y = split_download(ro, io, prec, H1 * H2)
    assert np.array_equal(y.reshape(H2, H1).T.reshape(-1),
                          x.astype(np.complex128).reshape(H1, H2).reshape(-1))
    p.destroy()
    # guru64 entry point, 2-D
    shape = (20, 30)
    dims, tot = row_major(shape)
    x = inputs_c(tot, 14, prec)
    ri, ii = split_upload(x, prec)
    ro, io = fa.DeviceBuffer(ri.nbytes), fa.DeviceBuffer(ri.nbytes)
    p = fa.plan_guru_split_dft(prec, dims, [], ri.ptr, ii.ptr, ro.ptr, io.ptr, wide=True)
    p.execute()
    y = split_download(ro, io, prec, tot)
    ref = oracle.fftn(x.astype(np.complex128).reshape(shape), -1).reshape(-1)
    assert oracle.rel_l2(y, ref) < TOL[prec]
    p.destroy()
