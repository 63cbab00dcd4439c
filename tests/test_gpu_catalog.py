"""Every kernel of the catalog against the oracle.

The planner picks among the catalog's geometries and variants (threads per
line, lines per workgroup, radix order, radix 8 or 16) by timing them under
FFTW_MEASURE, so any entry can end up in a plan.  Each entry is forced into
a plan (fftw_amd_force_kernel) for a problem that routes to its length,
mapping and IO kind, and the result is compared with the oracle on a batch
with a ragged last workgroup.  The reference verifies every plan its planner
evaluates the same way, against its own transform (tests/hook.c:195-253).

Problems by IO kind (H lines, L the kernel length):
  plain row   c2c, n = L, contiguous lines
  plain col   c2c, n = L, lines adjacent (stride H, distance 1)
  tout        c2c, n = L, output transposed (stride H, distance 1)
  r2c / c2r   n = 2L, contiguous or strided real lines (alternating)
  r2c-r2r     R2HC, DHT, REDFT10, RODFT10 (n = 2L), REDFT00 (n = L+1),
              RODFT00 (n = L-1)
  c2r-r2r     HC2R, REDFT01, RODFT01 (n = 2L)
  dct4        REDFT11, RODFT11 (n = 2L)
  r2c-full / c2r-full   odd n = L
  bluestein   c2c of the largest prime p with 2p - 1 <= L
Column kernels of the real kinds run on strided lines (stride H).
"""
import numpy as np
import pytest

import fftw3_amd as fa
import oracle
from gpu_util import TOL, cdt, rdt

pytestmark = pytest.mark.gpu

try:
    fa.lib()
    CAT = fa.catalog()
except Exception:  # library not built: collection still works
    CAT = []

F = {"R2HC": fa.FFTW_R2HC, "HC2R": fa.FFTW_HC2R, "DHT": fa.FFTW_DHT, "REDFT00": fa.FFTW_REDFT00,
     "REDFT01": fa.FFTW_REDFT01, "REDFT10": fa.FFTW_REDFT10, "REDFT11": fa.FFTW_REDFT11,
     "RODFT00": fa.FFTW_RODFT00, "RODFT01": fa.FFTW_RODFT01, "RODFT10": fa.FFTW_RODFT10,
     "RODFT11": fa.FFTW_RODFT11}


def lines_for(e):
    """Two full workgroups plus a ragged one, bounded in size."""
    H = 2 * e["nl"] + 1
    while H > 3 and e["L"] * H > (1 << 21):
        H = max(3, H // 2)
    return H


def key_of(e):
    return "stockham %s L=%d rad=%s %s %s tpl=%d nl=%d " % (
        "f64" if e["prec"] == "double" else "f32", e["L"], "x".join(map(str, e["rad"])), e["mode"],
        e["io"], e["tpl"], e["nl"])


def largest_prime(limit):
    def isp(n):
        return n > 1 and all(n % d for d in range(2, int(n ** 0.5) + 1))
    p = limit
    while not isp(p):
        p -= 1
    return p


def seed_of(e):
    return 1000 + e["index"]


def check_uses(p, e):
    info = p.pass_info()
    assert any(s.startswith(key_of(e)) for s in info), (key_of(e), info)


def run_c2c(e, dims, hdims, span_in, span_out, sign, n_check):
    prec = e["prec"]
    x = (oracle.uniform(2 * span_in, seed_of(e)).astype(rdt(prec))).view(cdt(prec))
    bi = fa.DeviceBuffer.from_array(x)
    bo = fa.DeviceBuffer(span_out * np.dtype(cdt(prec)).itemsize)
    p = fa.plan_guru_dft(prec, dims, hdims, bi.ptr, bo.ptr, sign)
    check_uses(p, e)
    p.execute()
    y = bo.download(cdt(prec), span_out)
    ref = np.zeros(span_out, np.complex128)
    oracle.guru_dft(dims, hdims, x.astype(np.complex128), ref, sign)
    p.destroy()
    return y, ref


def run_r2c(e, n, H, strided):
    prec = e["prec"]
    nh = n // 2 + 1
    if strided:
        dims, hdims, span_in = [(n, 2, 1)], [(H, 2 * n, nh)], 2 * n * H
    else:
        dims, hdims, span_in = [(n, 1, 1)], [(H, n, nh)], n * H
    x = oracle.uniform(span_in, seed_of(e)).astype(rdt(prec))
    bi = fa.DeviceBuffer.from_array(x)
    bo = fa.DeviceBuffer(nh * H * np.dtype(cdt(prec)).itemsize)
    p = fa.plan_guru_dft_r2c(prec, dims, hdims, bi.ptr, bo.ptr)
    check_uses(p, e)
    p.execute()
    y = bo.download(cdt(prec), nh * H)
    ref = np.zeros(nh * H, np.complex128)
    oracle.guru_r2c(dims, hdims, x.astype(np.float64), ref)
    p.destroy()
    return y, ref


def run_c2r(e, n, H, strided):
    prec = e["prec"]
    nh = n // 2 + 1
    if strided:
        dims, hdims, span_out = [(n, 1, 2)], [(H, nh, 2 * n)], 2 * n * H
    else:
        dims, hdims, span_out = [(n, 1, 1)], [(H, nh, n)], n * H
    x = (oracle.uniform(2 * nh * H, seed_of(e)).astype(rdt(prec))).view(cdt(prec))
    bi = fa.DeviceBuffer.from_array(x)
    bo = fa.DeviceBuffer(span_out * np.dtype(rdt(prec)).itemsize)
    bo.upload(np.zeros(span_out, rdt(prec)))
    p = fa.plan_guru_dft_c2r(prec, dims, hdims, bi.ptr, bo.ptr)
    check_uses(p, e)
    p.execute()
    y = bo.download(rdt(prec), span_out)
    ref = np.zeros(span_out, np.float64)
    oracle.guru_c2r(dims, hdims, x.astype(np.complex128), ref)
    p.destroy()
    return y, ref


def run_2d_real(e, n0, n1, H, forward):
    """2-D r2c (forward) or c2r of n0 x n1 reals, howmany H, out of place."""
    prec = e["prec"]
    nh = n1 // 2 + 1
    if forward:
        x = oracle.uniform(n0 * n1 * H, seed_of(e)).astype(rdt(prec))
        bi = fa.DeviceBuffer.from_array(x)
        bo = fa.DeviceBuffer(n0 * nh * H * np.dtype(cdt(prec)).itemsize)
        p = fa.plan_many_dft_r2c(prec, [n0, n1], H, bi.ptr, None, 1, n0 * n1, bo.ptr, None, 1, n0 * nh)
        check_uses(p, e)
        p.execute()
        y = bo.download(cdt(prec), n0 * nh * H)
        ref = oracle.rfftn(x.astype(np.float64).reshape(H, n0, n1), axes_batch=1).reshape(-1)
    else:
        # Hermitian input: the r2c of real data (the c2r reads only that half)
        xr = oracle.uniform(n0 * n1 * H, seed_of(e)).reshape(H, n0, n1)
        X = oracle.rfftn(xr, axes_batch=1)
        x = X.astype(cdt(prec)).reshape(-1)
        bi = fa.DeviceBuffer.from_array(x)
        bo = fa.DeviceBuffer(n0 * n1 * H * np.dtype(rdt(prec)).itemsize)
        p = fa.plan_many_dft_c2r(prec, [n0, n1], H, bi.ptr, None, 1, n0 * nh, bo.ptr, None, 1, n0 * n1)
        check_uses(p, e)
        p.execute()
        y = bo.download(rdt(prec), n0 * n1 * H)
        ref = oracle.irfftn(x.astype(np.complex128).reshape(H, n0, nh), n1, axes_batch=1).reshape(-1)
    p.destroy()
    return y, ref


def run_r2r(e, kind, n, H, col):
    prec = e["prec"]
    if col:
        dims, hdims = [(n, H, H)], [(H, 1, 1)]
    else:
        dims, hdims = [(n, 1, 1)], [(H, n, n)]
    span = n * H
    x = oracle.uniform(span, seed_of(e) + kind).astype(rdt(prec))
    bi = fa.DeviceBuffer.from_array(x)
    bo = fa.DeviceBuffer(x.nbytes)
    p = fa.plan_guru_r2r(prec, dims, hdims, bi.ptr, bo.ptr, [kind])
    check_uses(p, e)
    p.execute()
    y = bo.download(rdt(prec), span)
    ref = np.zeros(span, np.float64)
    oracle.guru_r2r(dims, hdims, [kind], x.astype(np.float64), ref)
    p.destroy()
    return y, ref


def cases(e):
    """(label, thunk) pairs for catalog entry e."""
    L, H, io, mode = e["L"], lines_for(e), e["io"], e["mode"]
    sign = -1 if e["index"] % 2 == 0 else 1
    strided = e["index"] % 2 == 1
    out = []
    if io == "plain" and mode == "row":
        out.append(("c2c", lambda: run_c2c(e, [(L, 1, 1)], [(H, L, L)], L * H, L * H, sign, L)))
    elif io == "plain":
        out.append(("c2c col", lambda: run_c2c(e, [(L, H, H)], [(H, 1, 1)], L * H, L * H, sign, L)))
    elif io == "tout":
        out.append(("c2c tout", lambda: run_c2c(e, [(L, 1, H)], [(H, L, 1)], L * H, L * H, sign, L)))
    elif io == "bluestein":
        p = largest_prime((L + 1) // 2)
        out.append(("bluestein n=%d" % p,
                    lambda: run_c2c(e, [(p, 1, 1)], [(H, p, p)], p * H, p * H, sign, p)))
    elif io == "r2c":
        out.append(("r2c", lambda: run_r2c(e, 2 * L, H, strided)))
    elif io == "c2r":
        out.append(("c2r", lambda: run_c2r(e, 2 * L, H, strided)))
    elif io == "r2c-full":
        out.append(("r2c full", lambda: run_r2c(e, L, H, strided)))
    elif io == "c2r-full":
        out.append(("c2r full", lambda: run_c2r(e, L, H, strided)))
    elif io in ("r2c-xr", "c2r-xr"):
        # 2-D problems whose first dim is nl * m (nl = the kernel's rows per
        # tile), howmany 2: the row pass runs the first radix-nl stage of
        # the column DFT
        out.append(("%s 2-D" % io, lambda: run_2d_real(e, e["nl"] * 4, 2 * L, 2, io == "r2c-xr")))
    else:
        col = mode == "col"
        if io == "r2c-r2r":
            kinds = [("R2HC", 2 * L), ("DHT", 2 * L), ("REDFT10", 2 * L), ("RODFT10", 2 * L),
                     ("REDFT00", L + 1)]
            if L - 1 >= 2:
                kinds.append(("RODFT00", L - 1))
        elif io == "c2r-r2r":
            kinds = [("HC2R", 2 * L), ("REDFT01", 2 * L), ("RODFT01", 2 * L)]
        else:
            kinds = [("REDFT11", 2 * L), ("RODFT11", 2 * L)]
        for name, n in kinds:
            out.append(("%s n=%d" % (name, n),
                        lambda name=name, n=n: run_r2r(e, F[name], n, H, col)))
    return out


@pytest.fixture(scope="module")
def forced(eng):
    yield
    fa.force_kernel(-1)


def _id(e):
    return "%d-%s-L%d-%s-%s-v%d" % (e["index"], "f64" if e["prec"] == "double" else "f32", e["L"],
                                   e["mode"], e["io"], e["variant"])


@pytest.mark.parametrize("e", CAT, ids=[_id(e) for e in CAT])
def test_catalog_entry(forced, e):
    fa.force_kernel(e["index"])
    try:
        for label, thunk in cases(e):
            y, ref = thunk()
            err = oracle.rel_l2(y, ref)
            assert err <= TOL[e["prec"]], (label, key_of(e), err)
    finally:
        fa.force_kernel(-1)


def test_catalog_is_complete():
    """Every entry has a routing above (no kind left unchecked)."""
    kinds = {(e["mode"], e["io"]) for e in CAT}
    known = {("row", "plain"), ("col", "plain"), ("row", "tout"), ("row", "bluestein"),
             ("row", "r2c"), ("row", "c2r"), ("row", "r2c-full"), ("row", "c2r-full"),
             ("row", "r2c-r2r"), ("col", "r2c-r2r"), ("row", "c2r-r2r"), ("col", "c2r-r2r"),
             ("row", "dct4"), ("col", "dct4"), ("row", "r2c-xr"), ("row", "c2r-xr")}
    assert CAT and kinds <= known, kinds - known
