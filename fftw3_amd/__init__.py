"""Python binding of libfftw3_amd.so (the MI355X FFT engine) via ctypes.

The product is the C-ABI library; this module mirrors its fftw3.h entry
points (same names, argument meaning, NULL-plan error behaviour) so tests and
the benchmark can drive it from numpy.  Arrays passed to plans are device
pointers (ints); DeviceBuffer moves numpy data to and from device memory.

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfftw3_amd.so")

FFTW_FORWARD = -1
FFTW_BACKWARD = 1
FFTW_MEASURE = 0
FFTW_DESTROY_INPUT = 1 << 0
FFTW_UNALIGNED = 1 << 1
FFTW_CONSERVE_MEMORY = 1 << 2
FFTW_EXHAUSTIVE = 1 << 3
FFTW_PRESERVE_INPUT = 1 << 4
FFTW_PATIENT = 1 << 5
FFTW_ESTIMATE = 1 << 6
FFTW_WISDOM_ONLY = 1 << 21
FFTW_AMD_MFMA = 1 << 24
# r2r kinds (api/fftw3.h:89-100)
FFTW_R2HC, FFTW_HC2R, FFTW_DHT = 0, 1, 2
FFTW_REDFT00, FFTW_REDFT01, FFTW_REDFT10, FFTW_REDFT11 = 3, 4, 5, 6
FFTW_RODFT00, FFTW_RODFT01, FFTW_RODFT10, FFTW_RODFT11 = 7, 8, 9, 10
FFTW_AMD_BLOCKED = 1 << 25

_lib = None


class IODim(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("is_", ctypes.c_int), ("os", ctypes.c_int)]


class Request(ctypes.Structure):
    """fftw_amd_request: kind 0 c2c / 1 r2c / 2 c2r, prec 0 double / 1 float."""
    _fields_ = [("kind", ctypes.c_int), ("prec", ctypes.c_int), ("sign", ctypes.c_int),
                ("rank", ctypes.c_int), ("n", ctypes.c_int * 3), ("howmany", ctypes.c_int),
                ("in_", ctypes.c_void_p)]


class Server:
    """fftw_amd_server handle (continuous-batching serving loop)."""

    def __init__(self, budget_bytes, max_rounds=4, per_round=64):
        self.h = lib().fftw_amd_server_create(int(budget_bytes), max_rounds, per_round)
        if not self.h:
            raise RuntimeError("server creation failed: " + last_error())

    @staticmethod
    def request(kind, prec, shape, howmany, inp, sign=-1):
        r = Request()
        r.kind = {"c2c": 0, "r2c": 1, "c2r": 2}[kind]
        r.prec = 0 if prec == "double" else 1
        r.sign = sign
        r.rank = len(shape)
        for i, v in enumerate(shape):
            r.n[i] = v
        r.howmany = howmany
        r.in_ = inp
        return r

    def prepare(self, r):
        return lib().fftw_amd_server_prepare(self.h, ctypes.byref(r))

    def submit(self, r):
        return lib().fftw_amd_server_submit(self.h, ctypes.byref(r))

    def poll(self, i):
        return lib().fftw_amd_server_poll(self.h, i)

    def wait(self, i):
        out = ctypes.c_void_p()
        lat = ctypes.c_double()
        rc = lib().fftw_amd_server_wait(self.h, i, ctypes.byref(out), ctypes.byref(lat))
        if rc:
            raise RuntimeError("request %d failed" % i)
        return out.value, lat.value

    def release(self, i):
        return lib().fftw_amd_server_release(self.h, i)

    def stats(self):
        p50, p99 = ctypes.c_double(), ctypes.c_double()
        done, tr = ctypes.c_longlong(), ctypes.c_longlong()
        el, used = ctypes.c_double(), ctypes.c_size_t()
        lib().fftw_amd_server_stats(self.h, ctypes.byref(p50), ctypes.byref(p99), ctypes.byref(done),
                                    ctypes.byref(tr), ctypes.byref(el), ctypes.byref(used))
        return {"p50_ms": p50.value, "p99_ms": p99.value, "completed": done.value,
                "transforms": tr.value, "elapsed_s": el.value, "hbm_in_use": used.value}

    def destroy(self):
        if self.h:
            lib().fftw_amd_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class IODim64(ctypes.Structure):
    _fields_ = [("n", ctypes.c_ssize_t), ("is_", ctypes.c_ssize_t), ("os", ctypes.c_ssize_t)]


def build(force=False):
    """Generate the kernel catalog and compile the library (make)."""
    subprocess.check_call(["python3", os.path.join(HERE, "gen_kernels.py")])
    cmd = ["make", "-C", HERE, "-j%d" % max(1, min(16, os.cpu_count() or 1))]
    if force:
        cmd.append("-B")
    subprocess.check_call(cmd)


def lib():
    """Load libfftw3_amd.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libfftw3_amd.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, u = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint
    ipp = ctypes.POINTER(ctypes.c_int)
    for P in ("fftw_", "fftwf_"):
        def f(name, res, args):
            fn = getattr(L, P + name)
            fn.restype = res
            fn.argtypes = args
        f("plan_many_dft", vp, [ip, ipp, ip, vp, ipp, ip, ip, vp, ipp, ip, ip, ip, u])
        f("plan_dft", vp, [ip, ipp, vp, vp, ip, u])
        f("plan_dft_1d", vp, [ip, vp, vp, ip, u])
        f("plan_dft_2d", vp, [ip, ip, vp, vp, ip, u])
        f("plan_dft_3d", vp, [ip, ip, ip, vp, vp, ip, u])
        f("plan_guru_dft", vp, [ip, ctypes.POINTER(IODim), ip, ctypes.POINTER(IODim), vp, vp, ip, u])
        f("plan_guru64_dft", vp, [ip, ctypes.POINTER(IODim64), ip, ctypes.POINTER(IODim64), vp, vp, ip, u])
        f("plan_many_dft_r2c", vp, [ip, ipp, ip, vp, ipp, ip, ip, vp, ipp, ip, ip, u])
        f("plan_many_dft_c2r", vp, [ip, ipp, ip, vp, ipp, ip, ip, vp, ipp, ip, ip, u])
        f("plan_dft_r2c", vp, [ip, ipp, vp, vp, u])
        f("plan_dft_c2r", vp, [ip, ipp, vp, vp, u])
        f("plan_dft_r2c_1d", vp, [ip, vp, vp, u])
        f("plan_dft_c2r_1d", vp, [ip, vp, vp, u])
        f("plan_dft_r2c_2d", vp, [ip, ip, vp, vp, u])
        f("plan_dft_c2r_2d", vp, [ip, ip, vp, vp, u])
        f("plan_dft_r2c_3d", vp, [ip, ip, ip, vp, vp, u])
        f("plan_dft_c2r_3d", vp, [ip, ip, ip, vp, vp, u])
        f("plan_guru_dft_r2c", vp, [ip, ctypes.POINTER(IODim), ip, ctypes.POINTER(IODim), vp, vp, u])
        f("plan_guru_dft_c2r", vp, [ip, ctypes.POINTER(IODim), ip, ctypes.POINTER(IODim), vp, vp, u])
        f("plan_guru64_dft_r2c", vp, [ip, ctypes.POINTER(IODim64), ip, ctypes.POINTER(IODim64), vp, vp, u])
        f("plan_guru64_dft_c2r", vp, [ip, ctypes.POINTER(IODim64), ip, ctypes.POINTER(IODim64), vp, vp, u])
        f("plan_guru_split_dft", vp, [ip, ctypes.POINTER(IODim), ip, ctypes.POINTER(IODim), vp, vp, vp, vp, u])
        f("plan_guru64_split_dft", vp, [ip, ctypes.POINTER(IODim64), ip, ctypes.POINTER(IODim64), vp, vp, vp, vp,
                                        u])
        f("plan_guru_split_dft_r2c", vp, [ip, ctypes.POINTER(IODim), ip, ctypes.POINTER(IODim), vp, vp, vp, u])
        f("plan_guru_split_dft_c2r", vp, [ip, ctypes.POINTER(IODim), ip, ctypes.POINTER(IODim), vp, vp, vp, u])
        f("plan_guru64_split_dft_r2c", vp, [ip, ctypes.POINTER(IODim64), ip, ctypes.POINTER(IODim64), vp, vp, vp,
                                            u])
        f("plan_guru64_split_dft_c2r", vp, [ip, ctypes.POINTER(IODim64), ip, ctypes.POINTER(IODim64), vp, vp, vp,
                                            u])
        f("execute_split_dft", None, [vp, vp, vp, vp, vp])
        f("execute_split_dft_r2c", None, [vp, vp, vp, vp])
        f("execute_split_dft_c2r", None, [vp, vp, vp, vp])
        f("plan_many_r2r", vp, [ip, ipp, ip, vp, ipp, ip, ip, vp, ipp, ip, ip, ipp, u])
        f("plan_r2r", vp, [ip, ipp, vp, vp, ipp, u])
        f("plan_r2r_1d", vp, [ip, vp, vp, ip, u])
        f("plan_r2r_2d", vp, [ip, ip, vp, vp, ip, ip, u])
        f("plan_r2r_3d", vp, [ip, ip, ip, vp, vp, ip, ip, ip, u])
        f("plan_guru_r2r", vp, [ip, ctypes.POINTER(IODim), ip, ctypes.POINTER(IODim), vp, vp, ipp, u])
        f("plan_guru64_r2r", vp, [ip, ctypes.POINTER(IODim64), ip, ctypes.POINTER(IODim64), vp, vp, ipp, u])
        f("execute_r2r", None, [vp, vp, vp])
        f("execute", None, [vp])
        f("execute_dft", None, [vp, vp, vp])
        f("execute_dft_r2c", None, [vp, vp, vp])
        f("execute_dft_c2r", None, [vp, vp, vp])
        f("destroy_plan", None, [vp])
        f("copy_plan", vp, [vp])
        f("cleanup", None, [])
        f("forget_wisdom", None, [])
        f("set_timelimit", None, [ctypes.c_double])
        f("export_wisdom_to_string", ctypes.c_void_p, [])
        f("import_wisdom_from_string", ip, [ctypes.c_char_p])
        f("sprint_plan", ctypes.c_void_p, [vp])
        f("malloc", vp, [ctypes.c_size_t])
        f("free", None, [vp])
        f("flops", None, [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double)])
        f("estimate_cost", ctypes.c_double, [vp])
        f("cost", ctypes.c_double, [vp])
        f("alignment_of", ip, [vp])
        f("init_threads", ip, [])
        f("plan_with_nthreads", None, [ip])
        f("planner_nthreads", ip, [])
    L.fftw_amd_init.restype = ctypes.c_int
    L.fftw_amd_last_error.restype = ctypes.c_char_p
    L.fftw_amd_set_stream.argtypes = [vp]
    L.fftw_amd_get_stream.restype = vp
    L.fftw_amd_plan_set_stream.argtypes = [vp, vp]
    L.fftw_amd_set_async.argtypes = [ip]
    for n in ("fftw_amd_memcpy_htod", "fftw_amd_memcpy_dtoh", "fftw_amd_memcpy_dtod"):
        getattr(L, n).argtypes = [vp, vp, ctypes.c_size_t]
        getattr(L, n).restype = ctypes.c_int
    L.fftw_amd_memset.argtypes = [vp, ip, ctypes.c_size_t]
    L.fftw_amd_stream_copy.argtypes = [vp, vp, ctypes.c_size_t, ip, ip, vp]
    L.fftw_amd_stream_copy.restype = ctypes.c_int
    L.fftw_amd_synchronize.restype = ctypes.c_int
    L.fftw_amd_mem_free.restype = ctypes.c_size_t
    for n in ("fftw_amd_fill_uniform", "fftw_amd_fill_gauss"):
        getattr(L, n).argtypes = [vp, ip, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_uint64]
        getattr(L, n).restype = ctypes.c_int
    L.fftw_amd_plan_npasses.argtypes = [vp]
    L.fftw_amd_plan_pass_info.argtypes = [vp, ip, ctypes.c_char_p, ip]
    L.fftw_amd_plan_bytes.argtypes = [vp]
    L.fftw_amd_plan_bytes.restype = ctypes.c_double
    L.fftw_amd_plan_pass_bytes.argtypes = [vp, ip]
    L.fftw_amd_plan_pass_bytes.restype = ctypes.c_double
    L.fftw_amd_plan_time.argtypes = [vp, vp, vp, ip, ctypes.POINTER(ctypes.c_double)]
    L.fftw_amd_plan_time.restype = ctypes.c_double
    L.fftw_amd_plan_batch_dft_1d.restype = vp
    L.fftw_amd_plan_batch_dft_1d.argtypes = [ip, ip, ipp, ipp, ctypes.POINTER(vp),
                                             ctypes.POINTER(vp), ip, u]
    for n_ in ("fftw_amd_batch_execute", "fftw_amd_batch_execute_direct"):
        getattr(L, n_).argtypes = [vp]
        getattr(L, n_).restype = ctypes.c_int
    L.fftw_amd_batch_info.argtypes = [vp, ctypes.c_char_p, ip]
    L.fftw_amd_batch_stream.argtypes = [vp]
    L.fftw_amd_batch_stream.restype = vp
    L.fftw_amd_batch_destroy.argtypes = [vp]
    L.fftw_amd_catalog_size.restype = ctypes.c_int
    L.fftw_amd_catalog_entry.argtypes = [ip, ipp]
    L.fftw_amd_catalog_entry.restype = ctypes.c_int
    L.fftw_amd_force_kernel.argtypes = [ip]
    L.fftw_amd_force_kernel.restype = ctypes.c_int
    L.fftw_amd_plan_use_mfma.argtypes = [vp]
    L.fftw_amd_mfma_probe.argtypes = [vp]
    L.fftw_amd_server_create.restype = vp
    L.fftw_amd_server_create.argtypes = [ctypes.c_size_t, ip, ip]
    L.fftw_amd_server_prepare.argtypes = [vp, ctypes.POINTER(Request)]
    L.fftw_amd_server_submit.argtypes = [vp, ctypes.POINTER(Request)]
    L.fftw_amd_server_submit.restype = ctypes.c_longlong
    L.fftw_amd_server_poll.argtypes = [vp, ctypes.c_longlong]
    L.fftw_amd_server_wait.argtypes = [vp, ctypes.c_longlong, ctypes.POINTER(vp),
                                       ctypes.POINTER(ctypes.c_double)]
    L.fftw_amd_server_release.argtypes = [vp, ctypes.c_longlong]
    L.fftw_amd_server_stats.argtypes = [vp] + [ctypes.POINTER(ctypes.c_double)] * 2 + \
        [ctypes.POINTER(ctypes.c_longlong)] * 2 + [ctypes.POINTER(ctypes.c_double),
                                                   ctypes.POINTER(ctypes.c_size_t)]
    L.fftw_amd_server_destroy.argtypes = [vp]
    L.fftw_amd_event_create.restype = vp
    L.fftw_amd_event_record.argtypes = [vp, vp]
    L.fftw_amd_event_elapsed.argtypes = [vp, vp]
    L.fftw_amd_event_elapsed.restype = ctypes.c_double
    L.fftw_amd_event_destroy.argtypes = [vp]
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [vp]
    L._libc = libc
    _lib = L
    return L


def prefix(prec):
    return "fftw_" if prec == "double" else "fftwf_"


def fn(prec, name):
    return getattr(lib(), prefix(prec) + name)


def last_error():
    return lib().fftw_amd_last_error().decode()


def init():
    r = lib().fftw_amd_init()
    if r != 0:
        raise RuntimeError("fftw_amd_init failed: " + last_error())


class DeviceBuffer:
    """Device allocation from fftw_malloc (device memory)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = lib().fftw_malloc(max(1, self.nbytes))
        if not self.ptr:
            raise MemoryError("fftw_malloc(%d) failed" % self.nbytes)

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        r = lib().fftw_amd_memcpy_htod(self.ptr + offset, a.ctypes.data, a.nbytes)
        if r:
            raise RuntimeError("htod copy failed (%d)" % r)

    def download(self, dtype, count=None, offset=0):
        dt = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dt.itemsize
        out = np.empty(count, dtype=dt)
        r = lib().fftw_amd_memcpy_dtoh(out.ctypes.data, self.ptr + offset, out.nbytes)
        if r:
            raise RuntimeError("dtoh copy failed (%d)" % r)
        return out

    def fill_uniform(self, prec, count, seed, stream=0):
        r = lib().fftw_amd_fill_uniform(self.ptr, 0 if prec == "double" else 1, count, seed, stream)
        if r:
            raise RuntimeError("fill failed (%d)" % r)

    def fill_gauss(self, prec, count, seed, stream=0):
        r = lib().fftw_amd_fill_gauss(self.ptr, 0 if prec == "double" else 1, count, seed, stream)
        if r:
            raise RuntimeError("fill failed (%d)" % r)

    def free(self):
        if self.ptr:
            lib().fftw_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Plan:
    """Owns an fftw_plan / fftwf_plan handle."""

    def __init__(self, handle, prec):
        if not handle:
            raise RuntimeError("planner returned NULL: " + last_error())
        self.h = handle
        self.prec = prec

    def execute(self):
        fn(self.prec, "execute")(self.h)

    def execute_dft(self, i, o):
        fn(self.prec, "execute_dft")(self.h, i, o)

    def execute_r2c(self, i, o):
        fn(self.prec, "execute_dft_r2c")(self.h, i, o)

    def execute_c2r(self, i, o):
        fn(self.prec, "execute_dft_c2r")(self.h, i, o)

    def execute_r2r(self, i, o):
        fn(self.prec, "execute_r2r")(self.h, i, o)

    def execute_split(self, ri, ii, ro, io):
        fn(self.prec, "execute_split_dft")(self.h, ri, ii, ro, io)

    def execute_split_r2c(self, i, ro, io):
        fn(self.prec, "execute_split_dft_r2c")(self.h, i, ro, io)

    def execute_split_c2r(self, ri, ii, o):
        fn(self.prec, "execute_split_dft_c2r")(self.h, ri, ii, o)

    def npasses(self):
        return lib().fftw_amd_plan_npasses(self.h)

    def pass_info(self):
        out = []
        for i in range(self.npasses()):
            buf = ctypes.create_string_buffer(512)
            lib().fftw_amd_plan_pass_info(self.h, i, buf, 512)
            out.append(buf.value.decode())
        return out

    def bytes(self):
        return lib().fftw_amd_plan_bytes(self.h)

    def pass_bytes(self):
        return [lib().fftw_amd_plan_pass_bytes(self.h, i) for i in range(self.npasses())]

    def time(self, i, o, iters=10):
        n = self.npasses()
        per = (ctypes.c_double * max(1, n))()
        tot = lib().fftw_amd_plan_time(self.h, i, o, iters, per)
        if tot < 0:
            raise RuntimeError("plan timing failed: " + last_error())
        return tot, [per[k] for k in range(n)]

    def sprint(self):
        p = fn(self.prec, "sprint_plan")(self.h)
        s = ctypes.cast(p, ctypes.c_char_p).value.decode()
        lib()._libc.free(p)
        return s

    def destroy(self):
        if self.h:
            fn(self.prec, "destroy_plan")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Batch:
    """fftw_amd_plan_batch_dft_1d handle: groups of (n, howmany, in, out)."""

    def __init__(self, prec, groups, sign=FFTW_FORWARD, flags=FFTW_ESTIMATE):
        k = len(groups)
        n = (ctypes.c_int * k)(*[g[0] for g in groups])
        h = (ctypes.c_int * k)(*[g[1] for g in groups])
        ins = (ctypes.c_void_p * k)(*[g[2] for g in groups])
        outs = (ctypes.c_void_p * k)(*[g[3] for g in groups])
        self.h = lib().fftw_amd_plan_batch_dft_1d(0 if prec == "double" else 1, k, n, h, ins, outs,
                                                  sign, flags)
        if not self.h:
            raise RuntimeError("batch planner returned NULL: " + last_error())

    def execute(self, graph=True):
        f = lib().fftw_amd_batch_execute if graph else lib().fftw_amd_batch_execute_direct
        r = f(self.h)
        if r:
            raise RuntimeError("batch execute failed (%d): %s" % (r, last_error()))

    def info(self):
        buf = ctypes.create_string_buffer(1024)
        lib().fftw_amd_batch_info(self.h, buf, 1024)
        return buf.value.decode()

    def stream(self):
        return lib().fftw_amd_batch_stream(self.h)

    def destroy(self):
        if self.h:
            lib().fftw_amd_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


IO_NAMES = ["plain", "tout", "r2c", "c2r", "bluestein", "r2c-r2r", "c2r-r2r", "dct4", "r2c-full",
            "c2r-full", "r2c-xr", "c2r-xr"]


def catalog():
    """The kernel catalog: one dict per entry (fftw_amd_catalog_entry)."""
    L = lib()
    out = []
    info = (ctypes.c_int * 22)()
    for i in range(L.fftw_amd_catalog_size()):
        if L.fftw_amd_catalog_entry(i, info):
            raise RuntimeError("catalog entry %d" % i)
        v = list(info)
        out.append(dict(index=i, prec="double" if v[0] == 0 else "float", L=v[1],
                        mode="row" if v[2] == 0 else "col", io=IO_NAMES[v[3]], tpl=v[4], nl=v[5],
                        variant=v[6], rad=v[8:8 + v[7]], lds=v[20], loop=v[21]))
    return out


def force_kernel(i):
    """Plans created from now on use catalog entry i where the planner picks
    a kernel of its precision, length, mapping and IO kind (-1: normal)."""
    rc = lib().fftw_amd_force_kernel(i)
    if rc:
        raise ValueError("catalog index %d out of range" % i)


def _iarr(v):
    if v is None:
        return None
    return (ctypes.c_int * len(v))(*v)


def plan_many_dft(prec, n, howmany, inp, inembed, istride, idist, out, onembed, ostride, odist,
                  sign, flags=FFTW_ESTIMATE):
    h = fn(prec, "plan_many_dft")(len(n), _iarr(n), howmany, inp, _iarr(inembed), istride, idist,
                                  out, _iarr(onembed), ostride, odist, sign, flags)
    return Plan(h, prec)


def plan_many_dft_r2c(prec, n, howmany, inp, inembed, istride, idist, out, onembed, ostride, odist,
                      flags=FFTW_ESTIMATE):
    h = fn(prec, "plan_many_dft_r2c")(len(n), _iarr(n), howmany, inp, _iarr(inembed), istride, idist,
                                      out, _iarr(onembed), ostride, odist, flags)
    return Plan(h, prec)


def plan_many_dft_c2r(prec, n, howmany, inp, inembed, istride, idist, out, onembed, ostride, odist,
                      flags=FFTW_ESTIMATE):
    h = fn(prec, "plan_many_dft_c2r")(len(n), _iarr(n), howmany, inp, _iarr(inembed), istride, idist,
                                      out, _iarr(onembed), ostride, odist, flags)
    return Plan(h, prec)


def plan_many_r2r(prec, n, howmany, inp, inembed, istride, idist, out, onembed, ostride, odist, kinds,
                  flags=FFTW_ESTIMATE):
    h = fn(prec, "plan_many_r2r")(len(n), _iarr(n), howmany, inp, _iarr(inembed), istride, idist,
                                  out, _iarr(onembed), ostride, odist, _iarr(kinds), flags)
    return Plan(h, prec)


def plan_guru_r2r(prec, dims, hdims, inp, out, kinds, flags=FFTW_ESTIMATE, wide=False):
    name = "plan_guru64_r2r" if wide else "plan_guru_r2r"
    h = fn(prec, name)(len(dims), _dims(dims, wide), len(hdims), _dims(hdims, wide), inp, out,
                       _iarr(kinds), flags)
    return Plan(h, prec)


def _dims(dims, wide):
    T = IODim64 if wide else IODim
    arr = (T * max(1, len(dims)))()
    for i, (n, s_in, s_out) in enumerate(dims):
        arr[i].n, arr[i].is_, arr[i].os = n, s_in, s_out
    return arr


def plan_guru_dft(prec, dims, hdims, inp, out, sign, flags=FFTW_ESTIMATE, wide=False):
    name = "plan_guru64_dft" if wide else "plan_guru_dft"
    h = fn(prec, name)(len(dims), _dims(dims, wide), len(hdims), _dims(hdims, wide), inp, out, sign, flags)
    return Plan(h, prec)


def plan_guru_dft_r2c(prec, dims, hdims, inp, out, flags=FFTW_ESTIMATE, wide=False):
    name = "plan_guru64_dft_r2c" if wide else "plan_guru_dft_r2c"
    h = fn(prec, name)(len(dims), _dims(dims, wide), len(hdims), _dims(hdims, wide), inp, out, flags)
    return Plan(h, prec)


def plan_guru_dft_c2r(prec, dims, hdims, inp, out, flags=FFTW_ESTIMATE, wide=False):
    name = "plan_guru64_dft_c2r" if wide else "plan_guru_dft_c2r"
    h = fn(prec, name)(len(dims), _dims(dims, wide), len(hdims), _dims(hdims, wide), inp, out, flags)
    return Plan(h, prec)


def plan_guru_split_dft(prec, dims, hdims, ri, ii, ro, io, flags=FFTW_ESTIMATE, wide=False):
    """Split-complex guru c2c; strides count reals.  Forward on (ri, ii);
    pass (ii, ri, io, ro) for the backward transform."""
    name = "plan_guru64_split_dft" if wide else "plan_guru_split_dft"
    h = fn(prec, name)(len(dims), _dims(dims, wide), len(hdims), _dims(hdims, wide), ri, ii, ro, io, flags)
    return Plan(h, prec)


def plan_guru_split_dft_r2c(prec, dims, hdims, inp, ro, io, flags=FFTW_ESTIMATE, wide=False):
    name = "plan_guru64_split_dft_r2c" if wide else "plan_guru_split_dft_r2c"
    h = fn(prec, name)(len(dims), _dims(dims, wide), len(hdims), _dims(hdims, wide), inp, ro, io, flags)
    return Plan(h, prec)


def plan_guru_split_dft_c2r(prec, dims, hdims, ri, ii, out, flags=FFTW_ESTIMATE, wide=False):
    name = "plan_guru64_split_dft_c2r" if wide else "plan_guru_split_dft_c2r"
    h = fn(prec, name)(len(dims), _dims(dims, wide), len(hdims), _dims(hdims, wide), ri, ii, out, flags)
    return Plan(h, prec)


def cdtype(prec):
    return np.complex128 if prec == "double" else np.complex64


def rdtype(prec):
    return np.float64 if prec == "double" else np.float32
