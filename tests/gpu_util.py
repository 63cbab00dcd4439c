"""Helpers for GPU parity tests: run a transform through the C-ABI library
(libfftw3_amd.so via fftw3_amd) and the same transform through the oracle,
on the same seeded inputs."""
import numpy as np

import fftw3_amd as fa
import oracle

# parity bars (relative L2 against the long-double oracle)
TOL = {"double": 1e-13, "float": 2e-6}


def cdt(prec):
    return np.complex128 if prec == "double" else np.complex64


def rdt(prec):
    return np.float64 if prec == "double" else np.float32


def inputs_c(count, seed, prec):
    """count complex values from the counter-based stream, rounded to prec."""
    x = oracle.uniform(2 * count, seed).astype(rdt(prec))
    return x.view(cdt(prec))


def inputs_r(count, seed, prec):
    return oracle.uniform(count, seed).astype(rdt(prec))


def c2c_many(prec, n, howmany, sign, seed=1, inplace=False, flags=fa.FFTW_ESTIMATE):
    """plan_many_dft over contiguous batches; returns (gpu, oracle) outputs."""
    tot = int(np.prod(n)) * howmany
    x = inputs_c(tot, seed, prec)
    bi = fa.DeviceBuffer.from_array(x)
    bo = bi if inplace else fa.DeviceBuffer(x.nbytes)
    dist = int(np.prod(n))
    p = fa.plan_many_dft(prec, list(n), howmany, bi.ptr, None, 1, dist, bo.ptr, None, 1, dist, sign, flags)
    if not (flags & fa.FFTW_ESTIMATE):
        bi.upload(x)  # MEASURE may overwrite the arrays
    p.execute()
    y = bo.download(cdt(prec), tot)
    ref = oracle.fftn(x.astype(np.complex128).reshape((howmany,) + tuple(n)), sign,
                      axes_batch=1).reshape(-1)
    p.destroy()
    return y, ref, x


def guru_c2c(prec, dims, hdims, sign, span_in, span_out, seed=1, inplace=False):
    """Arbitrary guru layout on flat buffers of span_in / span_out complex."""
    x = inputs_c(span_in, seed, prec)
    bi = fa.DeviceBuffer.from_array(x)
    if inplace:
        bo = bi
        span_out = span_in
    else:
        bo = fa.DeviceBuffer(span_out * np.dtype(cdt(prec)).itemsize)
        bo.upload(np.zeros(span_out, cdt(prec)))
    p = fa.plan_guru_dft(prec, dims, hdims, bi.ptr, bo.ptr, sign)
    p.execute()
    y = bo.download(cdt(prec), span_out)
    ref = np.zeros(span_out, np.complex128) if not inplace else x.astype(np.complex128).copy()
    src = x.astype(np.complex128)
    if inplace:
        oracle.guru_dft(dims, hdims, ref, ref, sign)
    else:
        oracle.guru_dft(dims, hdims, src, ref, sign)
    p.destroy()
    return y, ref


def mask_for(dims, hdims, span, use_os=True):
    """Boolean mask of the output positions a guru problem writes."""
    m = np.zeros(span, bool)
    idx = np.zeros(1, np.int64)
    for (n, s_in, s_out) in list(hdims) + list(dims):
        s = s_out if use_os else s_in
        idx = (idx[:, None] + s * np.arange(n)[None, :]).reshape(-1)
    m[idx] = True
    return m
