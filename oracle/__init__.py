"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this module, and only to check results.
The product library (fftw3_amd) never imports or links it.

The oracle restates FFTW 3.3's DFT semantics from the mathematics; see
oracle/oracle.h for the reference files it follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

LD = 0  # long-double internal precision (accuracy oracle)
D = 1   # double internal precision (CPU throughput baseline)


class _IODim(ctypes.Structure):
    _fields_ = [("n", ctypes.c_long), ("is_", ctypes.c_long), ("os", ctypes.c_long)]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(_IODim)
        for name in ("orc_dft",):
            f = getattr(L, name)
            f.argtypes = [ctypes.c_int, ctypes.c_int, ip, ctypes.c_int, ip, dp, dp, ctypes.c_int]
            f.restype = ctypes.c_int
        for name in ("orc_r2c", "orc_c2r"):
            f = getattr(L, name)
            f.argtypes = [ctypes.c_int, ctypes.c_int, ip, ctypes.c_int, ip, dp, dp]
            f.restype = ctypes.c_int
        L.orc_r2r.argtypes = [ctypes.c_int, ctypes.c_int, ip, ctypes.c_int, ip,
                              ctypes.POINTER(ctypes.c_int), dp, dp]
        L.orc_r2r.restype = ctypes.c_int
        L.orc_dft_direct.argtypes = [ctypes.c_long, ctypes.c_int, dp, dp]
        L.orc_dft_bins.argtypes = [ctypes.c_long, ctypes.c_int, dp,
                                   ctypes.POINTER(ctypes.c_long), ctypes.c_long, dp]
        L.orc_cexp.argtypes = [ctypes.c_long, ctypes.c_long, ctypes.c_int, dp]
        L.orc_splitmix64.argtypes = [ctypes.c_uint64]
        L.orc_splitmix64.restype = ctypes.c_uint64
        L.orc_fill_uniform.argtypes = [dp, ctypes.c_long, ctypes.c_uint64, ctypes.c_uint64]
        L.orc_fill_gauss.argtypes = [dp, ctypes.c_long, ctypes.c_uint64, ctypes.c_uint64]
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _dims(dims):
    arr = (_IODim * max(1, len(dims)))()
    for i, (n, is_, os_) in enumerate(dims):
        arr[i].n, arr[i].is_, arr[i].os = n, is_, os_
    return arr


def num_threads():
    return lib().orc_num_threads()


def guru_dft(dims, hdims, x, out, sign, prec=LD):
    """Complex DFT on flat complex128 arrays with FFTW guru strides (complex units)."""
    assert x.dtype == np.complex128 and out.dtype == np.complex128
    r = lib().orc_dft(prec, len(dims), _dims(dims), len(hdims), _dims(hdims),
                      _dp(x), _dp(out), sign)
    if r != 0:
        raise ValueError("oracle rejected problem")
    return out


def guru_r2c(dims, hdims, x, out, prec=LD):
    assert x.dtype == np.float64 and out.dtype == np.complex128
    r = lib().orc_r2c(prec, len(dims), _dims(dims), len(hdims), _dims(hdims), _dp(x), _dp(out))
    if r != 0:
        raise ValueError("oracle rejected problem")
    return out


def guru_c2r(dims, hdims, x, out, prec=LD):
    assert x.dtype == np.complex128 and out.dtype == np.float64
    r = lib().orc_c2r(prec, len(dims), _dims(dims), len(hdims), _dims(hdims), _dp(x), _dp(out))
    if r != 0:
        raise ValueError("oracle rejected problem")
    return out


def _row_major_dims(shape, elem_mult_in=1, elem_mult_out=1, shape_in=None, shape_out=None):
    shape_in = shape_in or shape
    shape_out = shape_out or shape
    dims = []
    si = so = 1
    for d in reversed(range(len(shape))):
        dims.append((shape[d], si, so))
        si *= shape_in[d]
        so *= shape_out[d]
    return list(reversed(dims)), si, so


def fftn(x, sign=-1, axes_batch=0, prec=LD):
    """DFT over the trailing x.ndim - axes_batch axes of a C-contiguous array."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    shape = x.shape[axes_batch:]
    dims, tot, _ = _row_major_dims(shape)
    hdims = []
    s = tot
    for d in reversed(range(axes_batch)):
        hdims.append((x.shape[d], s, s))
        s *= x.shape[d]
    hdims = list(reversed(hdims))
    out = np.empty_like(x)
    return guru_dft(dims, hdims, x.reshape(-1), out.reshape(-1), sign, prec).reshape(x.shape)


def rfftn(x, axes_batch=0, prec=LD):
    x = np.ascontiguousarray(x, dtype=np.float64)
    shape = list(x.shape[axes_batch:])
    oshape = shape[:-1] + [shape[-1] // 2 + 1]
    dims, ti, to = _row_major_dims(shape, shape_out=oshape)
    hdims = []
    si, so = ti, to
    for d in reversed(range(axes_batch)):
        hdims.append((x.shape[d], si, so))
        si *= x.shape[d]
        so *= x.shape[d]
    out = np.empty(tuple(x.shape[:axes_batch]) + tuple(oshape), dtype=np.complex128)
    guru_r2c(dims, list(reversed(hdims)), x.reshape(-1), out.reshape(-1), prec)
    return out


def irfftn(X, n_last, axes_batch=0, prec=LD):
    """Unnormalized c2r; X has last dim n_last//2+1."""
    X = np.ascontiguousarray(X, dtype=np.complex128)
    ishape = list(X.shape[axes_batch:])
    shape = ishape[:-1] + [n_last]
    dims, ti, to = _row_major_dims(shape, shape_in=ishape)
    hdims = []
    si, so = ti, to
    for d in reversed(range(axes_batch)):
        hdims.append((X.shape[d], si, so))
        si *= X.shape[d]
        so *= X.shape[d]
    out = np.empty(tuple(X.shape[:axes_batch]) + tuple(shape), dtype=np.float64)
    guru_c2r(dims, list(reversed(hdims)), X.reshape(-1), out.reshape(-1), prec)
    return out


def guru_r2r(dims, hdims, kinds, x, out, prec=LD):
    """Real-to-real guru transform; kinds are fftw_r2r_kind values."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    k = (ctypes.c_int * max(1, len(kinds)))(*kinds)
    rc = lib().orc_r2r(prec, len(dims), _dims(dims), len(hdims), _dims(hdims), k, _dp(x), _dp(out))
    if rc:
        raise ValueError("orc_r2r rejected the problem")
    return out


def r2rn(x, kinds, axes_batch=0, prec=LD):
    """Separable r2r over the trailing len(kinds) axes of a C-contiguous
    array (leading axes_batch axes are a batch)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    shape = x.shape[axes_batch:]
    dims, s = [], 1
    for n in reversed(shape):
        dims.insert(0, (n, s, s))
        s *= n
    hd, hs = [], s
    for n in reversed(x.shape[:axes_batch]):
        hd.insert(0, (n, hs, hs))
        hs *= n
    out = np.empty_like(x)
    return guru_r2r(dims, hd, list(kinds), x, out, prec)


def dft_direct(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().orc_dft_direct(x.size, sign, _dp(x), _dp(out))
    return out


def dft_bins(x, bins, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    b = np.ascontiguousarray(bins, dtype=np.int64)
    out = np.empty(b.size, dtype=np.complex128)
    lib().orc_dft_bins(x.size, sign, _dp(x), b.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                       b.size, _dp(out))
    return out


def dft_bins_nd(x, bins, sign=-1):
    """Sampled bins of the DFT over all axes of a C-contiguous array
    (complex, or real with zero imaginary parts); bins: (nbins, ndim)."""
    real = not np.iscomplexobj(x)
    x = np.ascontiguousarray(x, dtype=np.float64 if real else np.complex128)
    b = np.ascontiguousarray(np.asarray(bins).reshape(-1, x.ndim), dtype=np.int64)
    n = (ctypes.c_long * x.ndim)(*x.shape)
    out = np.empty(b.shape[0], dtype=np.complex128)
    f = lib().orc_dft_bins_nd
    f.restype = None
    f.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_long), ctypes.c_int,
                  ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.POINTER(ctypes.c_long),
                  ctypes.c_long, ctypes.POINTER(ctypes.c_double)]
    f(x.ndim, n, sign, _dp(x), int(real), b.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
      b.shape[0], _dp(out))
    return out


def cexp(m, n, sign=-1):
    out = np.empty(2)
    lib().orc_cexp(m, n, sign, _dp(out))
    return complex(out[0], out[1])


def splitmix64(key):
    return lib().orc_splitmix64(key)


def uniform(count, seed, stream=0):
    """count reals 2u-1 from the counter-based stream."""
    x = np.empty(count, dtype=np.float64)
    lib().orc_fill_uniform(_dp(x), count, seed, stream)
    return x


def uniform_complex(count, seed, stream=0):
    return uniform(2 * count, seed, stream).view(np.complex128)


def gauss(count, seed, stream=0):
    x = np.empty(count, dtype=np.float64)
    lib().orc_fill_gauss(_dp(x), count, seed, stream)
    return x


def rel_l2(a, b):
    """||a - b||_2 / ||b||_2 with b the reference."""
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


def rel_linf(a, b):
    """max|a - b| / max|b| (libbench2 aerror convention, libbench2/verify-lib.c:36-58)."""
    a = np.asarray(a)
    b = np.asarray(b)
    mb = np.max(np.abs(b)) if b.size else 0.0
    return float(np.max(np.abs(a - b)) / (mb if mb > 0 else 1.0)) if a.size else 0.0
