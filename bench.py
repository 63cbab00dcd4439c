#!/usr/bin/env python3
"""Benchmark of the MI355X FFT engine on one GPU.

Headline workload (one "step"): the multidimensional/real configuration
  * 3-D complex-to-complex FP64, 512x512x512 in place, forward then backward;
  * 2-D real-to-complex FP32 2048x2048, howmany=64, out of place, then the
    matching complex-to-real;
inputs uniform in [-1, 1) from the counter-based generator (seed 2),
generated on the device; every array resident in HBM before timing; plans
made with FFTW_MEASURE (device autotuning).

value = algorithmic HBM bytes per step / seconds per step (GB/s), where
algorithmic bytes read every input once and write every output once:
c2c 2*N*sizeof(C) per direction; r2c/c2r N*sizeof(R) + N/n*(n/2+1)*sizeof(C).
gflops uses 5 N log2 N (c2c) and 2.5 N log2 N (r2c/c2r) per transform.

Before timing, the forward outputs of the timed plans are compared with
scipy.fft (pocketfft, FP64) on the same inputs, whole arrays for the 3-D
transform and two whole batches of the 2-D one, and the run fails above
rel-L2 1e-13 (FP64) / 2e-6 (FP32); the inverse transforms are checked by the
round trip.

The same JSON line carries the other legs of the metric, each bounded in
time: single-size batched c2c (4096 x 1024 FP64, hot and with buffers
rotated beyond the 256 MB memory-side cache; FP32 and FP64 sizes at batches
of 512 MiB to 1 GiB per array), the uneven-size batch
(per-call vs per-size vs persistent+hipGraph transforms/s), the serving loop
(p50/p99 latency and FFTs/s at a stated Poisson rate), the precision study
(FP64, FP32, FP16x2 matrix-core at 2^24 x 8) and a DCT-II sub-measure.

roofline: the dominant kernel (largest share of step time, per-pass HIP
event timing on the launch stream) -> its algorithmic bytes per launch /
its mean launch duration, against 8 TB/s, and against the streaming-copy
ceiling measured with the engine's own copy kernel.

cpu_baseline: FFTW's threaded CPU path when the host has a system
libfftw3; otherwise that is recorded as unavailable and scipy.fft with all
host cores is timed as a labelled stand-in (libbench2-style: repeat, keep the
minimum).
"""
import argparse
import ctypes
import ctypes.util
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8000.0  # GB/s, MI355X HBM3E
TOL = {"double": 1e-13, "float": 2e-6}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-legs", action="store_true",
                    help="headline step only (no batching/serving/precision/single-size legs)")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the comparison of the timed plans with scipy.fft (profiling runs)")
    ap.add_argument("--n3", type=int, default=512)
    ap.add_argument("--n2", type=int, default=2048)
    ap.add_argument("--howmany2", type=int, default=64)
    ap.add_argument("--estimate", action="store_true",
                    help="plan with FFTW_ESTIMATE instead of FFTW_MEASURE (device autotuning)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 2 GiB rank-0 copy plan (known traffic) for counter calibration")
    ap.add_argument("--workload", default="multid", choices=["multid", "mixed", "serve", "precision", "single"],
                    help="multid: the default bench line (all legs); the others run one leg alone")
    ap.add_argument("--requests", type=int, default=1500)
    ap.add_argument("--load", type=float, default=0.7, help="offered load as a fraction of capacity")
    return ap.parse_args()


def counter_uniform(count, seed, stream=0):
    """The device fill's input stream: x_i = 2u - 1, u = top 53 bits of
    splitmix64(seed ^ (stream << 40) ^ i) / 2^53."""
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) ^ (np.uint64(stream) << np.uint64(40))) ^ np.arange(count, dtype=np.uint64)
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return 2.0 * ((z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)) - 1.0


def rel_l2(a, b):
    a = np.asarray(a).ravel()
    b = np.asarray(b).ravel()
    num = np.linalg.norm(a.astype(np.complex128) - b.astype(np.complex128)) if a.size < (1 << 24) else \
        math.sqrt(sum(float(np.sum(np.abs(a[i:i + (1 << 24)].astype(np.complex128) -
                                          b[i:i + (1 << 24)].astype(np.complex128)) ** 2))
                      for i in range(0, a.size, 1 << 24)))
    den = math.sqrt(sum(float(np.sum(np.abs(b[i:i + (1 << 24)]) ** 2)) for i in range(0, b.size, 1 << 24)))
    return float(num / den)


def host_cores():
    """Cores this process may use (the box's CPU share), at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


class Timer:
    """HIP events on a stream (the stream the work is launched on)."""

    def __init__(self, L, stream=None):
        self.L, self.stream = L, stream
        self.e0, self.e1 = L.fftw_amd_event_create(), L.fftw_amd_event_create()

    def best(self, fn, reps, inner=1):
        fn()
        self.L.fftw_amd_synchronize()
        best = 1e30
        for _ in range(reps):
            self.L.fftw_amd_event_record(self.e0, self.stream)
            for _ in range(inner):
                fn()
            self.L.fftw_amd_event_record(self.e1, self.stream)
            self.L.fftw_amd_synchronize()
            best = min(best, self.L.fftw_amd_event_elapsed(self.e0, self.e1) / inner)
        return best

    def close(self):
        self.L.fftw_amd_event_destroy(self.e0)
        self.L.fftw_amd_event_destroy(self.e1)


# ---------------------------------------------------------------------------
# ceilings

def copy_ceiling(fa, L, nbytes=1 << 30):
    """Streaming-copy bandwidth: hipMemcpy device to device and the engine's
    own copy kernel (16-byte accesses, plain and non-temporal, several grid
    sizes), 1 GiB read + 1 GiB written, best of 5 each."""
    a = fa.DeviceBuffer(nbytes)
    b = fa.DeviceBuffer(nbytes)
    a.fill_uniform("float", nbytes // 4, 9, 0)
    t = Timer(L)
    res = {}
    ms = t.best(lambda: L.fftw_amd_memcpy_dtod(b.ptr, a.ptr, nbytes), 5)
    res["hipMemcpy"] = 2.0 * nbytes / (ms * 1e-3) / 1e9
    best_k = None
    for mode in (0, 1):
        for ctas in (1024, 2048, 4096, 8192):
            ms = t.best(lambda: L.fftw_amd_stream_copy(b.ptr, a.ptr, nbytes, mode, ctas, None), 5)
            gbs = 2.0 * nbytes / (ms * 1e-3) / 1e9
            res["kernel mode=%d ctas=%d" % (mode, ctas)] = round(gbs, 1)
            if best_k is None or gbs > best_k[0]:
                best_k = (gbs, mode, ctas)
    ok = np.array_equal(a.download(np.float32, 1 << 20), b.download(np.float32, 1 << 20))
    t.close()
    del a, b
    res["hipMemcpy"] = round(res["hipMemcpy"], 1)
    return {"GB_s": round(max(best_k[0], res["hipMemcpy"]), 1),
            "kernel_GB_s": round(best_k[0], 1), "kernel_mode": best_k[1], "kernel_ctas": best_k[2],
            "hipMemcpy_GB_s": res["hipMemcpy"], "copy_verified": bool(ok), "all": res}


# ---------------------------------------------------------------------------
# single-size batched c2c (the >= 70 % roofline configuration)

def single_size_c2c(fa, L, n=4096, howmany=1024, iters=50, rotate=8):
    """1-D c2c FP64, N=4096, howmany=1024, contiguous, out of place, seed 1,
    FFTW_MEASURE.  'hot': `iters` back-to-back executes on one buffer pair
    (134 MB, which the 256 MB memory-side cache can hold); 'rotated': the
    same plan executed round-robin over `rotate` buffer pairs (2 GiB in all),
    so no execute finds its data in that cache.  Checked in full against
    scipy.fft (FP64)."""
    import scipy.fft
    tot = n * howmany
    pairs = [(fa.DeviceBuffer(tot * 16), fa.DeviceBuffer(tot * 16)) for _ in range(rotate)]
    a, b = pairs[0]
    p = fa.plan_many_dft("double", [n], howmany, a.ptr, None, 1, n, b.ptr, None, 1, n,
                         fa.FFTW_FORWARD, fa.FFTW_MEASURE)
    for i, (x, _) in enumerate(pairs):
        x.fill_uniform("double", 2 * tot, 1, i)
    p.execute()
    L.fftw_amd_synchronize()
    x = a.download(np.complex128).reshape(howmany, n)
    err = rel_l2(b.download(np.complex128), scipy.fft.fft(x, axis=1, workers=host_cores()))
    t = Timer(L)
    ms_hot = t.best(lambda: p.execute(), 3, inner=iters)

    def rot():
        for (x_, y_) in pairs:
            p.execute_dft(x_.ptr, y_.ptr)

    ms_rot = t.best(rot, 3, inner=max(1, iters // rotate)) / rotate
    # per-pass event timing (kernel only, no inter-launch gaps)
    tk, _ = p.time(a.ptr, b.ptr, iters=20)
    t.close()
    by = 2.0 * tot * 16
    rec = {"ms_hot": round(ms_hot, 5), "GB_s_hot": round(by / (ms_hot * 1e-3) / 1e9, 1),
           "frac_hot": round(by / (ms_hot * 1e-3) / 1e9 / PEAK_HBM, 4),
           "ms_rotated": round(ms_rot, 5), "GB_s_rotated": round(by / (ms_rot * 1e-3) / 1e9, 1),
           "frac_rotated": round(by / (ms_rot * 1e-3) / 1e9 / PEAK_HBM, 4),
           "ms_kernel_event": round(tk, 5),
           "rel_l2_vs_scipy_fp64": err, "passes": p.pass_info()}
    p.destroy()
    del pairs
    return rec


def single_size_sweep(fa, L, iters=10):
    """Single-size batched c2c, out of place, FFTW_MEASURE, at batches whose
    arrays (512 MiB to 1 GiB each) exceed the 256 MB memory-side cache:
    per-execute time (best of 3 x `iters`), algorithmic bytes 2*N*H*sizeof(C)
    against 8 TB/s, and the first and last line checked against scipy.fft."""
    import scipy.fft
    out = {}
    t = Timer(L)
    for prec, n, H in (("float", 256, 262144), ("float", 1024, 65536), ("float", 4096, 16384),
                       ("double", 512, 131072), ("double", 4096, 16384)):
        cs = 16 if prec == "double" else 8
        dt = np.complex128 if prec == "double" else np.complex64
        tot = n * H
        a, b = fa.DeviceBuffer(tot * cs), fa.DeviceBuffer(tot * cs)
        a.fill_uniform(prec, 2 * tot, 1, 7)
        p = fa.plan_many_dft(prec, [n], H, a.ptr, None, 1, n, b.ptr, None, 1, n, fa.FFTW_FORWARD,
                             fa.FFTW_MEASURE)
        p.execute()
        L.fftw_amd_synchronize()
        x = a.download(dt, tot).reshape(H, n)[[0, H - 1]]
        y = b.download(dt, tot).reshape(H, n)[[0, H - 1]]
        err = rel_l2(y, scipy.fft.fft(x.astype(np.complex128), axis=1))
        ms = t.best(lambda: p.execute(), 3, inner=iters)
        by = 2.0 * tot * cs
        out["%s %d x%d" % ("fp64" if prec == "double" else "fp32", n, H)] = {
            "ms": round(ms, 5), "GB_s": round(by / (ms * 1e-3) / 1e9, 1),
            "frac": round(by / (ms * 1e-3) / 1e9 / PEAK_HBM, 4), "rel_l2_vs_scipy_fp64": err,
            "passes": [s.split(" lines=")[0] for s in p.pass_info()]}
        p.destroy()
        del a, b
    t.close()
    return out


def r2r_dct2(fa, L, shape=(2048, 2048), howmany=16, iters=20):
    """REDFT10 (DCT-II) FP32 over every dim of shape x howmany, out of place
    (FFTW_MEASURE): event-timed mean, algorithmic bytes 2*N*howmany*4, and a
    check of the first and last transforms against scipy.fft.dctn (type 2,
    norm None: FFTW's REDFT10 definition) on the same inputs."""
    import scipy.fft
    N = int(np.prod(shape))
    tot = N * howmany
    a = fa.DeviceBuffer(tot * 4)
    b = fa.DeviceBuffer(tot * 4)
    kinds = [fa.FFTW_REDFT10] * len(shape)
    p = fa.plan_many_r2r("float", list(shape), howmany, a.ptr, None, 1, N, b.ptr, None, 1, N, kinds,
                         fa.FFTW_MEASURE)
    a.fill_uniform("float", tot, 2, 0)
    p.execute()
    L.fftw_amd_synchronize()
    err = 0.0
    for h in (0, howmany - 1):
        x = a.download(np.float32, N, offset=h * N * 4).astype(np.float64).reshape(shape)
        err = max(err, rel_l2(b.download(np.float32, N, offset=h * N * 4),
                              scipy.fft.dctn(x, type=2, workers=host_cores()).reshape(-1)))
    t = Timer(L)
    ms = t.best(lambda: p.execute(), 3, inner=iters)
    t.close()
    gbs = 2.0 * tot * 4 / (ms * 1e-3) / 1e9
    rec = {"ms": round(ms, 5), "GB_s": round(gbs, 1), "frac": round(gbs / PEAK_HBM, 4),
           "rel_l2_vs_scipy_fp64": err, "passes": p.pass_info()}
    p.destroy()
    return rec


# ---------------------------------------------------------------------------
# precision study (C4)

def run_precision(fa, L, n=1 << 24, H=8):
    """c2c N=2^24, howmany=8, Gaussian N(0,1) (seed 4), FFTW_MEASURE: FP64,
    FP32 (FP64-generated twiddles) and FP32 with FP16 hi/lo matrix-core
    radix-16 passes (FFTW_AMD_MFMA).  Errors: every transform against
    scipy.fft in FP64 (FP32 variants), and transforms 0 and 7 against
    numpy.fft in long double (all variants)."""
    import scipy.fft
    a = fa.DeviceBuffer(n * H * 16)
    a.fill_gauss("double", 2 * n * H, 4, 0)
    x64 = a.download(np.complex128).reshape(H, n)
    t = time.time()
    ref64 = scipy.fft.fft(x64, axis=1, workers=host_cores())
    refld = {h: np.fft.fft(x64[h].astype(np.clongdouble)) for h in (0, H - 1)}
    log("precision references: %.1f s" % (time.time() - t))
    out = {}
    tm = Timer(L)
    for name, prec, mf in (("fp64", "double", False), ("fp32", "float", False),
                           ("fp32_fp16x2_mfma", "float", True)):
        cs = 16 if prec == "double" else 8
        if prec == "double":
            src = a
        else:
            src = fa.DeviceBuffer(n * H * 8)
            src.upload(x64.astype(np.complex64))
        b = fa.DeviceBuffer(n * H * cs)
        fl = fa.FFTW_MEASURE | (fa.FFTW_AMD_MFMA if mf else 0)
        p = fa.plan_many_dft(prec, [n], H, src.ptr, None, 1, n, b.ptr, None, 1, n, -1, fl)
        if prec == "double":
            a.fill_gauss("double", 2 * n * H, 4, 0)   # MEASURE may overwrite the arrays
        else:
            src.upload(x64.astype(np.complex64))
        p.execute()
        L.fftw_amd_synchronize()
        y = b.download(np.complex128 if prec == "double" else np.complex64).reshape(H, n)
        e_ld = max(float(np.linalg.norm((y[h].astype(np.clongdouble) - refld[h])) /
                         np.linalg.norm(refld[h])) for h in refld)
        emax = max(float(np.max(np.abs(y[h].astype(np.clongdouble) - refld[h])) /
                         np.max(np.abs(refld[h]))) for h in refld)
        rec = {"rel_l2_vs_longdouble": e_ld, "max_err_rel_vs_longdouble": emax}
        if prec == "float":
            rec["rel_l2_vs_scipy_fp64_all"] = rel_l2(y, ref64)
        ms = tm.best(lambda: p.execute(), 3, inner=3)
        byts = 2.0 * n * H * cs
        rec.update({"ms": round(ms, 4), "GB_s": round(byts / (ms * 1e-3) / 1e9, 1),
                    "passes": p.pass_info()})
        if mf:
            # matrix-core flops: 3 products of a 32x32 matrix by 32 reals per
            # radix-16 butterfly, n*H/16 butterflies per radix-16 stage
            mpass = [s_ for s_ in p.pass_info() if s_.startswith("mfma16")]
            stages = sum(len(s_.split("rad=")[1].split()[0].split("x")) for s_ in mpass)
            per_stage = 2.0 * (n * H / 16) * 3 * 32 * 32
            rec["mfma_passes"] = len(mpass)
            rec["mfma_dense_tflops"] = round(per_stage * stages / (ms * 1e-3) / 1e12, 2)
            rec["mfma_fraction_of_2.5PF_fp16"] = round(per_stage * stages / (ms * 1e-3) / 2.5e15, 4)
        out[name] = rec
        log("%-18s rel_l2(ld) %.3e  max %.3e  %.3f ms  %.1f GB/s" % (name, e_ld, emax, ms, rec["GB_s"]))
        p.destroy()
        del b
        if prec == "float":
            del src
    tm.close()
    ok = (out["fp64"]["rel_l2_vs_longdouble"] <= TOL["double"] and
          out["fp32"]["rel_l2_vs_longdouble"] <= TOL["float"] and
          out["fp32_fp16x2_mfma"]["rel_l2_vs_longdouble"] <= TOL["float"])
    return {"workload": "c2c N=2^24 howmany=8 Gaussian seed 4 (FFTW_MEASURE)", "within_bars": ok,
            "bars": {"fp64": TOL["double"], "fp32": TOL["float"]}, "variants": out}


# ---------------------------------------------------------------------------
# serving loop (C5)

def serving_requests(count, seed=5):
    """Request stream: 70% 1-D c2c, 20% 2-D r2c, 10% 3-D c2c (64^3..256^3);
    N log-uniform in 2^8..2^22 with ~10% non-powers of two or primes; 80%
    FP32 / 20% FP64; howmany log-uniform in 1..256, capped at 2^24 points."""
    u = (counter_uniform(8 * count, seed, 11) + 1.0) / 2.0
    odd = [3, 5, 7, 9, 15]
    primes = [257, 1009, 4099, 16411, 65537, 262147]
    reqs = []
    for i in range(count):
        a, b, c, d, e, f, g, h = u[8 * i: 8 * i + 8]
        prec = "float" if b < 0.8 else "double"
        if a < 0.7:
            k = 8 + int(c * 15)
            n = 1 << k
            if d < 0.1:
                if e < 0.5:
                    m = odd[int(f * len(odd))]
                    n = m << max(0, k - 3)
                else:
                    n = primes[int(f * len(primes))]
            shape, kind = (n,), "c2c"
        elif a < 0.9:
            k = 8 + int(c * 15)
            shape, kind = (1 << (k // 2), 1 << (k - k // 2)), "r2c"
        else:
            s = [64, 128, 256][int(c * 3)]
            shape, kind = (s, s, s), "c2c"
        N = int(np.prod(shape))
        hm = int(math.exp(g * math.log(256.0)))
        hm = max(1, min(hm, (1 << 24) // N))
        reqs.append((kind, prec, shape, hm))
    return reqs


def run_serve(fa, L, count=1500, load=0.7):
    """Serving loop: capacity with everything queued at once, then Poisson
    arrivals at `load` x capacity; p50/p99 latency and FFTs/s."""
    import threading
    reqs = serving_requests(count)
    shapes = {}
    for kind, prec, shape, hm in reqs:
        key = (kind, prec, shape)
        shapes[key] = max(shapes.get(key, 0), hm)
    inbuf = {}
    for i, ((kind, prec, shape), hm) in enumerate(sorted(shapes.items(), key=str)):
        N = int(np.prod(shape)) * hm
        reals = 2 * N if kind == "c2c" else N
        b = fa.DeviceBuffer(reals * (8 if prec == "double" else 4))
        b.fill_uniform(prec, reals, 5, i)
        inbuf[(kind, prec, shape)] = b
    budget = 8 << 30
    srv = fa.Server(budget, max_rounds=4, per_round=32)
    t0 = time.time()
    reqobjs = [fa.Server.request(kind, prec, shape, hm, inbuf[(kind, prec, shape)].ptr)
               for kind, prec, shape, hm in reqs]
    for r in reqobjs:
        srv.prepare(r)
    log("serving: %d requests, %d shapes, planned in %.2fs" % (len(reqs), len(shapes), time.time() - t0))
    transforms = sum(hm for _, _, _, hm in reqs)

    t = time.time()
    ids = [srv.submit(r) for r in reqobjs]
    for i in ids:
        srv.wait(i)
        srv.release(i)
    cap_s = time.time() - t
    cap_rps = len(reqs) / cap_s
    log("capacity: %.3f s for %d requests (%.1f req/s, %.1f FFTs/s)" %
        (cap_s, len(reqs), cap_rps, transforms / cap_s))
    srv.destroy()
    lam = load * cap_rps
    srv = fa.Server(budget, max_rounds=4, per_round=32)
    for r in reqobjs:
        srv.prepare(r)
    gaps = -np.log(1.0 - (counter_uniform(len(reqs), 5, 13) + 1.0) / 2.0) / lam
    arrivals = np.cumsum(gaps)
    ids = []

    def consumer():
        k = 0
        while k < len(reqs):
            if k < len(ids):
                srv.wait(ids[k])
                srv.release(ids[k])
                k += 1
            else:
                time.sleep(0.0002)

    th = threading.Thread(target=consumer)
    th.start()
    start = time.time()
    for r, ta in zip(reqobjs, arrivals):
        delay = start + ta - time.time()
        if delay > 0:
            time.sleep(delay)
        ids.append(srv.submit(r))
    th.join()
    wall = time.time() - start
    st = srv.stats()
    srv.destroy()
    rec = {"workload": "Poisson arrivals; 70% 1-D c2c, 20% 2-D r2c, 10% 3-D c2c (64^3..256^3); "
                       "N 2^8..2^22, ~10% non-pow2/prime; 80% FP32 / 20% FP64; howmany 1..256; seed 5",
           "requests": len(reqs), "transforms": transforms,
           "offered_rate_req_s": round(lam, 1), "load_fraction": load,
           "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
           "ffts_per_s": round(st["transforms"] / st["elapsed_s"], 1),
           "requests_per_s": round(st["completed"] / st["elapsed_s"], 1),
           "capacity_req_s": round(cap_rps, 1), "capacity_ffts_per_s": round(transforms / cap_s, 1),
           "hbm_budget_bytes": budget, "rounds_in_flight": 4, "requests_per_round": 32,
           "wall_s": round(wall, 3)}
    log("serving: p50 %.3f ms  p99 %.3f ms  %.1f FFTs/s at %.1f req/s offered" %
        (st["p50_ms"], st["p99_ms"], rec["ffts_per_s"], lam))
    return rec


# ---------------------------------------------------------------------------
# uneven-size batching (C3)

def mixed_sizes(total_log2=28, seed=3):
    """Uneven-size set: N in {1000, 1001, 2187, 4093, 4096, 65536, 2^20},
    batch b_N = floor(exp(U(0, ln 4096))) from the counter-based stream,
    rescaled so that sum N*b_N ~ 2^total_log2 with b_N >= 1."""
    sizes = [1000, 1001, 2187, 4093, 4096, 65536, 1 << 20]
    u = (counter_uniform(len(sizes), seed, 7) + 1.0) / 2.0
    b = [int(math.exp(x * math.log(4096.0))) for x in u]
    tot = sum(n * k for n, k in zip(sizes, b))
    s = float(1 << total_log2) / tot
    b = [max(1, int(round(k * s))) for k in b]
    return list(zip(sizes, b))


def run_mixed(fa, L):
    """The uneven-size set three ways, all planned with FFTW_MEASURE: one
    plan call per transform (a howmany=1 plan per size), one batched plan
    per size, and the persistent work-table batch replayed through a
    hipGraph (FP32 c2c forward).  The first and last transform of every size
    are checked against scipy.fft (FP64)."""
    import scipy.fft
    L.fftw_amd_set_async(1)
    mix = mixed_sizes()
    ntr = sum(k for _, k in mix)
    pts = sum(n * k for n, k in mix)
    log("mixed set (N, howmany): %s; %d transforms, %d points" % (mix, ntr, pts))
    bufs = []
    for i, (n, k) in enumerate(mix):
        bi = fa.DeviceBuffer(n * k * 8)
        bo = fa.DeviceBuffer(n * k * 8)
        bufs.append((bi, bo))
    t0 = time.time()
    single = [fa.plan_many_dft("float", [n], 1, bi.ptr, None, 1, n, bo.ptr, None, 1, n,
                               fa.FFTW_FORWARD, fa.FFTW_MEASURE) for (n, k), (bi, bo) in zip(mix, bufs)]
    many = [fa.plan_many_dft("float", [n], k, bi.ptr, None, 1, n, bo.ptr, None, 1, n,
                             fa.FFTW_FORWARD, fa.FFTW_MEASURE) for (n, k), (bi, bo) in zip(mix, bufs)]
    groups = [(n, k, bi.ptr, bo.ptr) for (n, k), (bi, bo) in zip(mix, bufs)]
    bt = fa.Batch("float", groups, fa.FFTW_FORWARD, fa.FFTW_MEASURE)
    log("mixed: planned (MEASURE) in %.1f s; %s" % (time.time() - t0, bt.info()))
    for i, ((n, k), (bi, bo)) in enumerate(zip(mix, bufs)):
        bi.fill_uniform("float", 2 * n * k, 3, i)
    t = Timer(L)

    def per_call():
        for (n, k), (bi, bo), p in zip(mix, bufs, single):
            for j in range(k):
                p.execute_dft(bi.ptr + j * n * 8, bo.ptr + j * n * 8)

    t1 = t.best(per_call, 2)

    def per_size():
        for p in many:
            p.execute()

    t2 = t.best(per_size, 3)
    t.close()
    # the batch runs on its own stream: events on that stream
    tb = Timer(L, bt.stream())
    t3 = tb.best(lambda: bt.execute(), 5)
    tb.close()
    L.fftw_amd_synchronize()
    errs = {}
    for (n, k), (bi, bo) in zip(mix, bufs):
        e = 0.0
        for j in sorted({0, k - 1}):
            x = bi.download(np.complex64, n, offset=j * n * 8).astype(np.complex128)
            y = bo.download(np.complex64, n, offset=j * n * 8)
            e = max(e, rel_l2(y, scipy.fft.fft(x)))
        errs[str(n)] = e
    byts = 2.0 * pts * 8
    res = {}
    for name, tt in (("per_call", t1), ("per_size_batched", t2), ("persistent_graph", t3)):
        res[name] = {"ms": round(tt, 3), "transforms_per_s": round(ntr / (tt * 1e-3), 1),
                     "GB_s": round(byts / (tt * 1e-3) / 1e9, 1)}
        log("mixed %-18s %9.3f ms  %12.1f transforms/s  %8.1f GB/s" % (name, tt, ntr / (tt * 1e-3),
                                                                        byts / (tt * 1e-3) / 1e9))
    for p in single + many:
        p.destroy()
    bt.destroy()
    L.fftw_amd_set_async(1)
    return {"workload": "FP32 c2c N in {1000,1001,2187,4093,4096,65536,2^20}, seed 3, FFTW_MEASURE",
            "sizes_howmany": mix, "transforms": ntr, "points": pts,
            "speedup_persistent_vs_per_call": round(t1 / t3, 2),
            "speedup_persistent_vs_per_size": round(t2 / t3, 2), "modes": res,
            "rel_l2_first_last_vs_scipy_fp64": errs,
            "within_bar": all(v <= TOL["float"] for v in errs.values())}


# ---------------------------------------------------------------------------
# CPU baseline

def find_fftw():
    """System FFTW with its threads library, or None (ldconfig probe)."""
    names = {}
    for lib_ in ("fftw3", "fftw3_threads"):
        p = ctypes.util.find_library(lib_)
        if p:
            names[lib_] = p
    try:
        out = subprocess.run(["ldconfig", "-p"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        out = ""
    if "fftw3" in names and "fftw3_threads" in names:
        return names, out.count("libfftw3")
    return None, out.count("libfftw3")


def speed_min(fn, time_min=0.5, repeats=8, budget=12.0):
    """libbench2 speed() (libbench2/speed.c:58-88): double the iteration
    count until one timing reaches time_min, then keep the minimum over
    repeats, within a wall budget (fewer repeats when one run is long)."""
    it = 1
    while True:
        t = time.perf_counter()
        for _ in range(it):
            fn()
        dt = time.perf_counter() - t
        if dt >= time_min or it >= 1 << 20:
            break
        it *= 2
    best, done, t_all = dt / it, 1, dt
    while done < repeats and t_all + dt <= budget:
        t = time.perf_counter()
        for _ in range(it):
            fn()
        d = time.perf_counter() - t
        best = min(best, d / it)
        t_all += d
        done += 1
    return best, done, it


def fftw_reference_baseline(names, n3, n2, H2, cores):
    """FFTW's threaded path (threads/api.c:39-81): plan_with_nthreads(cores),
    FFTW_MEASURE, the headline step's transforms."""
    d = ctypes.CDLL(names["fftw3"], mode=ctypes.RTLD_GLOBAL)
    th = ctypes.CDLL(names["fftw3_threads"])
    th.fftw_init_threads()
    th.fftw_plan_with_nthreads(cores)
    vp = ctypes.c_void_p
    d.fftw_malloc.restype = vp
    d.fftw_plan_dft_3d.restype = vp
    d.fftw_plan_dft_3d.argtypes = [ctypes.c_int] * 3 + [vp, vp, ctypes.c_int, ctypes.c_uint]
    d.fftw_execute.argtypes = [vp]
    N = n3 ** 3
    buf = d.fftw_malloc(N * 16)
    pf = d.fftw_plan_dft_3d(n3, n3, n3, buf, buf, -1, 0)
    pb = d.fftw_plan_dft_3d(n3, n3, n3, buf, buf, 1, 0)
    arr = np.ctypeslib.as_array((ctypes.c_double * (2 * N)).from_address(buf))
    arr[:] = counter_uniform(2 * N, 2, 0)
    t, reps, it = speed_min(lambda: (d.fftw_execute(pf), d.fftw_execute(pb)), budget=20.0)
    by = 2 * 2.0 * N * 16
    return {"value": round(by / t / 1e9, 3), "unit": "GB/s", "cores": cores, "kind": "reference",
            "library": names, "sample": "system FFTW threaded, FFTW_MEASURE: 3-D c2c FP64 %d^3 "
                                        "in place fwd+bwd, min of %d x %d" % (n3, reps, it)}


def cpu_baseline(n3, n2, H2):
    """FFTW's threaded path when the host has it; otherwise recorded as
    unavailable, with scipy.fft (pocketfft) on all host cores as a labelled
    stand-in on the headline step's transforms: the 3-D FP64 transform at
    full size forward + backward, and the 2-D FP32 r2c + c2r on 8 of the 64
    batches (scaled to the step's bytes)."""
    import scipy.fft
    cores = host_cores()
    names, nlib = find_fftw()
    if names:
        try:
            return fftw_reference_baseline(names, n3, n2, H2, cores)
        except Exception as e:  # fall back to the stand-in, saying why
            log("system FFTW present but not usable: %s" % e)
    x = counter_uniform(2 * n3 ** 3, 2, 0).view(np.complex128).reshape(n3, n3, n3)
    t3, r3, i3 = speed_min(lambda: scipy.fft.ifftn(scipy.fft.fftn(x, workers=cores), workers=cores,
                                                   norm="forward"), budget=14.0)
    del x
    hs = 8
    xr = counter_uniform(hs * n2 * n2, 2, 1).astype(np.float32).reshape(hs, n2, n2)
    t2, r2, i2 = speed_min(lambda: scipy.fft.irfft2(scipy.fft.rfft2(xr, workers=cores), s=(n2, n2),
                                                    workers=cores), budget=8.0)
    del xr
    by3 = 2 * 2.0 * n3 ** 3 * 16
    by2 = 2 * H2 * (n2 * n2 * 4.0 + n2 * (n2 // 2 + 1) * 8.0)
    t_step = t3 + t2 * (H2 / hs)
    return {"value": round((by3 + by2) / t_step / 1e9, 3), "unit": "GB/s", "cores": cores,
            "kind": "stand-in",
            "fftw": "CPU FFTW baseline unavailable (no system libfftw3 / libfftw3_threads on the host; "
                    "ldconfig lists %d libfftw3 entries)" % nlib,
            "stand_in": "scipy.fft %s (pocketfft C++), workers=%d; not FFTW" % (scipy.__version__, cores),
            "ms_per_step": round(t_step * 1e3, 1),
            "sample": "headline step on the host: 3-D c2c FP64 %d^3 fwd+bwd at full size (min of %d x %d), "
                      "2-D r2c+c2r FP32 %dx%d on %d of %d batches scaled to %d (min of %d x %d); "
                      "libbench2-style timing" % (n3, r3, i3, n2, n2, hs, H2, H2, r2, i2)}


# ---------------------------------------------------------------------------
# headline

def check_timed_plans(fa, L, p3f, p3b, pr, pc, b3, rin, cmid, rout, n3, n2, H2):
    """Compare the timed plans with scipy.fft (FP64) on their own inputs:
    the whole 3-D forward transform, 2-D r2c batches 0 and H2-1, and both
    round trips in full.  Returns the errors; raises above the bars."""
    import scipy.fft
    cores = host_cores()
    tot3 = n3 ** 3
    nh = n2 // 2 + 1
    x3 = b3.download(np.complex128).reshape(n3, n3, n3)
    p3f.execute()
    L.fftw_amd_synchronize()
    X3 = b3.download(np.complex128).reshape(n3, n3, n3)
    e3 = rel_l2(X3, scipy.fft.fftn(x3, workers=cores))
    del X3
    p3b.execute()
    L.fftw_amd_synchronize()
    y3 = b3.download(np.complex128).reshape(n3, n3, n3)
    r3 = rel_l2(y3 / tot3, x3)
    del y3, x3
    pr.execute()
    L.fftw_amd_synchronize()
    # the multi-dimensional c2r destroys its input (as FFTW's does): read the
    # r2c output before running it
    e2 = 0.0
    for h in (0, H2 - 1):
        xr = rin.download(np.float32, n2 * n2, offset=h * n2 * n2 * 4).astype(np.float64).reshape(n2, n2)
        X = cmid.download(np.complex64, n2 * nh, offset=h * n2 * nh * 8)
        e2 = max(e2, rel_l2(X, scipy.fft.rfft2(xr, workers=cores)))
    pc.execute()
    L.fftw_amd_synchronize()
    xr = rin.download(np.float32).astype(np.float64)
    yr = rout.download(np.float32).astype(np.float64)
    r2 = rel_l2(yr / (n2 * n2), xr)
    del xr, yr
    errs = {"c2c3d_fwd_fp64_vs_scipy": e3, "c2c3d_roundtrip_fp64": r3,
            "r2c2d_fp32_vs_scipy_batches_0_last": e2, "r2c_c2r_roundtrip_fp32": r2}
    log("timed plans vs scipy.fft: %s" % ", ".join("%s %.3e" % kv for kv in errs.items()))
    bad = [k for k, v in errs.items() if not v <= (TOL["double"] if "fp64" in k else TOL["float"])]
    if bad:
        raise SystemExit("bench: timed plans fail the accuracy bar: %s" % errs)
    return errs


def kernel_symbol(info):
    """Demangled kernel name of a Stockham pass from its pass description
    ("stockham f64 L=512 rad=8x8x8 col plain tpl=64 nl=8")."""
    f = info.split()
    if f[0] != "stockham":
        return info
    kv = dict(x.split("=", 1) for x in f if "=" in x)
    rads = kv["rad"].split("x")
    real = "double" if f[1] == "f64" else "float"
    mode = 0 if f[4] == "row" else 1
    io = {"plain": 0, "tout": 1, "r2c": 2, "c2r": 3, "bluestein": 4, "r2c-r2r": 5, "c2r-r2r": 6, "dct4": 7,
          "r2c-full": 8, "c2r-full": 9}[f[5]]
    return "void fk::stockham<%s, %s, %d, %d, %s, fk::RadList<%s>, %s>(fk_args)" % (
        real, kv["nl"], mode, io, rads[0], ", ".join(rads), kv["tpl"])


def committed_traffic(symbol):
    """HBM bytes per dispatch of `symbol` from the committed FETCH_SIZE /
    WRITE_SIZE profile (profiles/traffic.json, calibrated on a copy kernel);
    None when that kernel was not profiled."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
        return t["kernels"].get(symbol)
    except (OSError, ValueError, KeyError):
        return None


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this engine has no multi-GPU path, and the "
              "distributed transposes are design only (requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import fftw3_amd as fa
    fa.init()
    L = fa.lib()
    if args.workload == "mixed":
        print(json.dumps(run_mixed(fa, L)), flush=True)
        return
    if args.workload == "serve":
        print(json.dumps(run_serve(fa, L, args.requests, args.load)), flush=True)
        return
    if args.workload == "precision":
        print(json.dumps(run_precision(fa, L)), flush=True)
        return
    # executes enqueue on the plan's stream without a host synchronize;
    # every timed region ends with fftw_amd_synchronize
    L.fftw_amd_set_async(1)
    if args.workload == "single":
        print(json.dumps(single_size_c2c(fa, L)), flush=True)
        return
    t_start = time.time()
    n3, n2, H2 = args.n3, args.n2, args.howmany2
    tot3 = n3 ** 3
    nh = n2 // 2 + 1

    # ---- buffers and inputs (device-resident) ----
    log("alloc 3d %d^3 fp64 (%.2f GiB), 2d r2c %dx%d x%d fp32" % (n3, tot3 * 16 / 2**30, n2, n2, H2))
    b3 = fa.DeviceBuffer(tot3 * 16)
    rin = fa.DeviceBuffer(H2 * n2 * n2 * 4)
    cmid = fa.DeviceBuffer(H2 * n2 * nh * 8)
    rout = fa.DeviceBuffer(H2 * n2 * n2 * 4)

    t0 = time.time()
    pflags = fa.FFTW_ESTIMATE if args.estimate else fa.FFTW_MEASURE
    p3f = fa.plan_many_dft("double", [n3] * 3, 1, b3.ptr, None, 1, tot3, b3.ptr, None, 1, tot3,
                           fa.FFTW_FORWARD, pflags)
    p3b = fa.plan_many_dft("double", [n3] * 3, 1, b3.ptr, None, 1, tot3, b3.ptr, None, 1, tot3,
                           fa.FFTW_BACKWARD, pflags)
    pr = fa.plan_many_dft_r2c("float", [n2, n2], H2, rin.ptr, None, 1, n2 * n2, cmid.ptr, None, 1,
                              n2 * nh, pflags)
    pc = fa.plan_many_dft_c2r("float", [n2, n2], H2, cmid.ptr, None, 1, n2 * nh, rout.ptr, None, 1,
                              n2 * n2, pflags)
    log("planned in %.2fs" % (time.time() - t0))
    # FFTW_MEASURE overwrote the arrays while timing candidates: fill now
    b3.fill_uniform("double", 2 * tot3, 2, 0)
    rin.fill_uniform("float", H2 * n2 * n2, 2, 1)
    plans = [("c2c3d_fwd", p3f), ("c2c3d_bwd", p3b), ("r2c2d", pr), ("c2r2d", pc)]
    for name, p in plans:
        for s in p.pass_info():
            log("  %s: %s" % (name, s))

    # ---- accuracy of the timed plans (untimed) ----
    errs = None
    if not args.no_check:
        t = time.time()
        errs = check_timed_plans(fa, L, p3f, p3b, pr, pc, b3, rin, cmid, rout, n3, n2, H2)
        log("checked in %.1f s" % (time.time() - t))
        b3.fill_uniform("double", 2 * tot3, 2, 0)

    if args.calib_copy:
        # plain copy of 2^27 complex FP64 (2 GiB read + 2 GiB write): the
        # reference dispatch for scaling FETCH_SIZE / WRITE_SIZE counters
        tmp = fa.DeviceBuffer(tot3 * 16)
        cp = fa.plan_guru_dft("double", [], [(tot3, 1, 1)], b3.ptr, tmp.ptr, fa.FFTW_FORWARD)
        cp.execute_dft(b3.ptr, tmp.ptr)
        cp.execute_dft(b3.ptr, tmp.ptr)
        L.fftw_amd_stream_copy(tmp.ptr, b3.ptr, tot3 * 16, 0, 0, None)
        L.fftw_amd_synchronize()
        log("calibration copy: 2 dispatches of %d bytes read + %d bytes written" % (tot3 * 16, tot3 * 16))
        del tmp, cp

    # ---- timed region: K steps, synchronized on both sides ----
    def step():
        p3f.execute()
        p3b.execute()
        pr.execute()
        pc.execute()

    for _ in range(args.warmup):
        step()
    L.fftw_amd_synchronize()
    e0, e1 = L.fftw_amd_event_create(), L.fftw_amd_event_create()
    L.fftw_amd_event_record(e0, None)
    for _ in range(args.steps):
        step()
    L.fftw_amd_event_record(e1, None)
    L.fftw_amd_synchronize()
    ms = L.fftw_amd_event_elapsed(e0, e1) / args.steps

    bytes3 = 2.0 * tot3 * 16
    bytes2 = H2 * (n2 * n2 * 4.0 + n2 * nh * 8.0)
    bytes_step = 2 * bytes3 + 2 * bytes2
    flops_step = 2 * 5.0 * tot3 * math.log2(tot3) + 2 * 2.5 * n2 * n2 * math.log2(n2 * n2) * H2
    value = bytes_step / (ms * 1e-3) / 1e9
    gflops = flops_step / (ms * 1e-3) / 1e9
    log("step %.3f ms  %.1f GB/s  %.1f GFLOP/s" % (ms, value, gflops))

    # ---- dominant kernel from per-pass event timing on the plan's stream ----
    groups = {}
    sub = {}
    for name, p in plans:
        ip, op = {"c2c3d_fwd": (b3, b3), "c2c3d_bwd": (b3, b3), "r2c2d": (rin, cmid),
                  "c2r2d": (cmid, rout)}[name]
        tot, per = p.time(ip.ptr, op.ptr, iters=5)
        sub[name] = tot
        for info, tp, by in zip(p.pass_info(), per, p.pass_bytes()):
            key = info.split(" lines=")[0]
            g = groups.setdefault(key, [0.0, 0.0, 0])
            g[0] += tp
            g[1] += by
            g[2] += 1
    key, (tms, tby, cnt) = max(groups.items(), key=lambda kv: kv[1][0])
    achieved = (tby / cnt) / (tms / cnt * 1e-3) / 1e9
    symbol = kernel_symbol(key)
    traffic = committed_traffic(symbol)
    log("dominant kernel: %s  launches/step=%d  mean %.3f ms  %.1f GB/s" % (key, cnt, tms / cnt, achieved))
    passes = {}
    for k, (tp, by, c) in sorted(groups.items(), key=lambda kv: -kv[1][0]):
        log("  %-90s %8.3f ms %7.1f GB/s x%d" % (k[:90], tp, by / (tp * 1e-3) / 1e9, c))
        passes[k] = {"ms_per_launch": round(tp / c, 4), "GB_s": round(by / (tp * 1e-3) / 1e9, 1),
                     "launches": c}
    del b3, rin, cmid, rout, p3f, p3b, pr, pc
    L.fftw_amd_synchronize()

    # ---- ceilings and the other legs of the metric ----
    legs = {}
    copy = copy_ceiling(fa, L)
    log("device copy: kernel %.1f GB/s (mode %d, %d WGs), hipMemcpy %.1f GB/s" %
        (copy["kernel_GB_s"], copy["kernel_mode"], copy["kernel_ctas"], copy["hipMemcpy_GB_s"]))
    if not args.no_legs:
        t = time.time()
        legs["single_size_c2c_fp64_4096x1024"] = c1 = single_size_c2c(fa, L)
        log("single-size c2c 4096 x 1024 fp64: hot %.4f ms (%.3f), rotated %.4f ms (%.3f), kernel %.4f ms"
            % (c1["ms_hot"], c1["frac_hot"], c1["ms_rotated"], c1["frac_rotated"], c1["ms_kernel_event"]))
        legs["single_size_c2c_large_batches"] = sw = single_size_sweep(fa, L)
        log("single-size c2c, large batches: " + ", ".join("%s %.3f" % (k, v["frac"]) for k, v in sw.items()))
        legs["r2r_redft10_fp32_2048x2048x16"] = r2r_dct2(fa, L)
        legs["mixed_batch_c3"] = run_mixed(fa, L)
        legs["serving_c5"] = run_serve(fa, L, args.requests, args.load)
        legs["precision_c4"] = run_precision(fa, L)
        log("legs in %.1f s" % (time.time() - t))

    cpu = None
    if not args.no_cpu:
        t = time.time()
        cpu = cpu_baseline(n3, n2, H2)
        log("cpu baseline (%s): %.3f GB/s on %d cores, %.1f s" % (cpu["kind"], cpu["value"], cpu["cores"],
                                                                 time.time() - t))

    config = {
        "workload": "3-D c2c FP64 %d^3 in-place fwd+bwd + 2-D r2c/c2r FP32 %dx%d x%d" % (n3, n2, n2, H2),
        "parallelism": "single GPU",
        "planner": "FFTW_ESTIMATE" if args.estimate else "FFTW_MEASURE (device autotune)",
        "gflops": round(gflops, 1),
        "bytes_per_step": bytes_step,
        "accuracy_of_timed_plans": errs,
        "sub_ms": {k: round(v, 4) for k, v in sub.items()},
        "passes": passes,
        "bench_wall_s": None,
    }
    config.update(legs)
    config["bench_wall_s"] = round(time.time() - t_start, 1)
    rec = {
        "metric": "HBM GB/s (algorithmic bytes) per batch, 1 GPU",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (3-D c2c) + f32 (2-D r2c/c2r)",
        "data": "synthetic: uniform [-1,1) counter-based splitmix64, seed 2, generated on device",
        "config": config,
        "roofline": {
            "bound": "hbm",
            "kernel": symbol,
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": tby / cnt,
            "mean_launch_ms": round(tms / cnt, 4),
            "device_copy_GBps": copy["GB_s"],
            "frac_of_copy": round(achieved / copy["GB_s"], 4),
            "copy": copy,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
