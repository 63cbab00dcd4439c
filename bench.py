#!/usr/bin/env python3
"""Benchmark of the MI355X FFT engine on one GPU.

Workload (one "step"): the multidimensional/real configuration
  * 3-D complex-to-complex FP64, 512x512x512 in place, forward then backward;
  * 2-D real-to-complex FP32 2048x2048, howmany=64, out of place, then the
    matching complex-to-real;
inputs uniform in [-1, 1) from the counter-based generator (seed 2),
generated on the device; every array resident in HBM before timing.

value = algorithmic HBM bytes per step / seconds per step (GB/s), where
algorithmic bytes read every input once and write every output once:
c2c 2*N*sizeof(C) per direction; r2c/c2r N*sizeof(R) + N/n*(n/2+1)*sizeof(C).
gflops uses 5 N log2 N (c2c) and 2.5 N log2 N (r2c/c2r) per transform.

roofline: the dominant kernel (largest share of step time, per-pass HIP
event timing on the launch stream) -> its algorithmic bytes per launch /
its mean launch duration, against 8 TB/s.  cpu_baseline: the oracle's
double-precision path on the host cores, on a bounded sample.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8000.0  # GB/s, MI355X HBM3E


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--n3", type=int, default=512)
    ap.add_argument("--n2", type=int, default=2048)
    ap.add_argument("--howmany2", type=int, default=64)
    ap.add_argument("--estimate", action="store_true",
                    help="plan with FFTW_ESTIMATE instead of FFTW_MEASURE (device autotuning)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 2 GiB rank-0 copy plan (known traffic) for counter calibration")
    ap.add_argument("--workload", default="multid", choices=["multid", "mixed", "serve", "precision"],
                    help="multid: the default bench line; mixed: uneven-size batching study; "
                         "serve: serving loop latency/throughput study")
    ap.add_argument("--requests", type=int, default=1500)
    ap.add_argument("--load", type=float, default=0.7, help="offered load as a fraction of capacity")
    return ap.parse_args()


def counter_uniform(count, seed, stream=0):
    """The device fill's input stream: x_i = 2u - 1, u = top 53 bits of
    splitmix64(seed ^ (stream << 40) ^ i) / 2^53."""
    import numpy as np
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) ^ (np.uint64(stream) << np.uint64(40))) ^ np.arange(count, dtype=np.uint64)
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return 2.0 * ((z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)) - 1.0


def rel_l2(a, b):
    import numpy as np
    a = np.asarray(a, dtype=np.complex128).ravel()
    b = np.asarray(b, dtype=np.complex128).ravel()
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


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


def run_precision(args):
    """Precision study: c2c N=2^24, howmany=8, Gaussian N(0,1) (seed 4):
    FP64, FP32 (FP64-generated twiddles) and FP32 with FP16 hi/lo
    matrix-core radix-16 passes, each against the long-double oracle."""
    import fftw3_amd as fa
    import oracle
    fa.init()
    L = fa.lib()
    n, H = 1 << 24, 8
    t = time.time()
    x64 = oracle.gauss(2 * n * H, 4).view(np.complex128)
    ref = oracle.fftn(x64.reshape(H, n), -1, axes_batch=1).reshape(-1)
    log("oracle (long double) %.1fs" % (time.time() - t))
    refmax = float(np.max(np.abs(ref)))
    e0, e1 = L.fftw_amd_event_create(), L.fftw_amd_event_create()
    out = {}
    for name, prec, mf in (("fp64", "double", False), ("fp32", "float", False),
                           ("fp32_fp16x2_mfma", "float", True)):
        x = x64 if prec == "double" else x64.astype(np.complex64)
        a = fa.DeviceBuffer.from_array(x)
        b = fa.DeviceBuffer(x.nbytes)
        fl = fa.FFTW_ESTIMATE | (fa.FFTW_AMD_MFMA if mf else 0)
        p = fa.plan_many_dft(prec, [n], H, a.ptr, None, 1, n, b.ptr, None, 1, n, -1, fl)
        sw = sum(1 for s_ in p.pass_info() if "tpl=512 nl=1" in s_ or "L=256" in s_) if mf else 0
        p.execute()
        y = b.download(x.dtype)
        err = oracle.rel_l2(y, ref)
        emax = float(np.max(np.abs(y.astype(np.complex128) - ref))) / refmax
        tot, per = p.time(a.ptr, b.ptr, iters=5)
        byts = 2.0 * n * H * np.dtype(x.dtype).itemsize
        rec = {"rel_l2": err, "max_err_rel": emax, "ms": round(tot, 4),
               "GB_s": round(byts / (tot * 1e-3) / 1e9, 1), "passes": p.pass_info()}
        if mf:
            # per 4096-point line: 3 stages x 256 butterflies x 3 split
            # products x (32 x 32 x 2) flops; two passes of n*H/4096 lines
            flops = 2.0 * (n * H / 4096) * 3 * 256 * 3 * 32 * 32 * 2
            rec["mfma_passes"] = sw
            rec["mfma_dense_tflops"] = round(flops / (tot * 1e-3) / 1e12, 2)
            rec["mfma_utilization_of_2.5PF"] = round(flops / (tot * 1e-3) / 2.5e15, 4)
        out[name] = rec
        log("%-18s rel_l2 %.3e  max %.3e  %.3f ms  %.1f GB/s" % (name, err, emax, tot, rec["GB_s"]))
        del a, b, p
    print(json.dumps({"metric": "precision study rel-L2 vs long-double oracle",
                      "config": {"workload": "c2c N=2^24 howmany=8 Gaussian seed 4"},
                      "variants": out}), flush=True)


def run_serve(args):
    """Serving loop: capacity with everything queued at once, then Poisson
    arrivals at `--load` x capacity; reports p50/p99 latency and FFTs/s."""
    import threading
    import fftw3_amd as fa
    fa.init()
    reqs = serving_requests(args.requests)
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
    reqobjs = []
    for kind, prec, shape, hm in reqs:
        r = fa.Server.request(kind, prec, shape, hm, inbuf[(kind, prec, shape)].ptr)
        reqobjs.append(r)
    for r in reqobjs:
        srv.prepare(r)
    log("serving: %d requests, %d shapes, planned in %.2fs" % (len(reqs), len(shapes), time.time() - t0))
    transforms = sum(hm for _, _, _, hm in reqs)

    def drain(ids):
        for i in ids:
            srv.wait(i)
            srv.release(i)

    # capacity: everything submitted at once
    t = time.time()
    ids = [srv.submit(r) for r in reqobjs]
    drain(ids)
    cap_s = time.time() - t
    cap_rps = len(reqs) / cap_s
    log("capacity: %.3f s for %d requests (%.1f req/s, %.1f FFTs/s)" %
        (cap_s, len(reqs), cap_rps, transforms / cap_s))
    srv.destroy()
    # Poisson arrivals at the offered rate
    lam = args.load * cap_rps
    srv = fa.Server(budget, max_rounds=4, per_round=32)
    for r in reqobjs:
        srv.prepare(r)
    gaps = -np.log(1.0 - (counter_uniform(len(reqs), 5, 13) + 1.0) / 2.0) / lam
    arrivals = np.cumsum(gaps)
    ids = []
    done = threading.Event()

    def consumer():
        k = 0
        while k < len(reqs):
            if k < len(ids):
                srv.wait(ids[k])
                srv.release(ids[k])
                k += 1
            else:
                time.sleep(0.0002)
        done.set()

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
    rec = {"metric": "served FFTs/s at p99 latency (mixed serving stream, 1 GPU)",
           "value": round(st["transforms"] / st["elapsed_s"], 1), "unit": "FFTs/s",
           "n_gpus": 1, "higher_is_better": True, "dtype": "f32 (80%) / f64 (20%)",
           "data": "synthetic seed 5",
           "config": {"workload": "serving mix: 70% 1-D c2c, 20% 2-D r2c, 10% 3-D c2c; N 2^8..2^22",
                      "requests": len(reqs), "transforms": transforms, "offered_rate_req_s": round(lam, 1),
                      "load_fraction": args.load, "hbm_budget_bytes": budget, "rounds_in_flight": 4,
                      "requests_per_round": 32},
           "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3),
           "requests_per_s": round(st["completed"] / st["elapsed_s"], 1),
           "capacity_req_s": round(cap_rps, 1), "wall_s": round(wall, 3)}
    log("p50 %.3f ms  p99 %.3f ms  %.1f FFTs/s at %.1f req/s offered" %
        (st["p50_ms"], st["p99_ms"], rec["value"], lam))
    print(json.dumps(rec), flush=True)


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


def run_mixed(args):
    """Throughput of the uneven-size set three ways: one plan call per
    transform, one batched plan per size, and the persistent work-table
    batch replayed through a hipGraph (FP32 c2c forward)."""
    import fftw3_amd as fa
    fa.init()
    L = fa.lib()
    L.fftw_amd_set_async(1)
    mix = mixed_sizes()
    ntr = sum(k for _, k in mix)
    pts = sum(n * k for n, k in mix)
    log("mixed set (N, howmany): %s; %d transforms, %d points" % (mix, ntr, pts))
    bufs = []
    for i, (n, k) in enumerate(mix):
        bi = fa.DeviceBuffer(n * k * 8)
        bi.fill_uniform("float", 2 * n * k, 3, i)
        bo = fa.DeviceBuffer(n * k * 8)
        bufs.append((bi, bo))
    e0, e1 = L.fftw_amd_event_create(), L.fftw_amd_event_create()

    def timed(fn, reps):
        fn()
        L.fftw_amd_synchronize()
        best = 1e30
        for _ in range(reps):
            L.fftw_amd_event_record(e0, None)
            fn()
            L.fftw_amd_event_record(e1, None)
            L.fftw_amd_synchronize()
            best = min(best, L.fftw_amd_event_elapsed(e0, e1))
        return best

    # (1) one plan per size with howmany=1, one execute per transform
    single = []
    for (n, k), (bi, bo) in zip(mix, bufs):
        single.append(fa.plan_many_dft("float", [n], 1, bi.ptr, None, 1, n, bo.ptr, None, 1, n,
                                       fa.FFTW_FORWARD, fa.FFTW_ESTIMATE))

    def per_call():
        for (n, k), (bi, bo), p in zip(mix, bufs, single):
            for t in range(k):
                p.execute_dft(bi.ptr + t * n * 8, bo.ptr + t * n * 8)

    t1 = timed(per_call, 2)
    # (2) one batched plan per size
    many = [fa.plan_many_dft("float", [n], k, bi.ptr, None, 1, n, bo.ptr, None, 1, n,
                             fa.FFTW_FORWARD, fa.FFTW_ESTIMATE) for (n, k), (bi, bo) in zip(mix, bufs)]

    def per_size():
        for p in many:
            p.execute()

    t2 = timed(per_size, 3)
    # (3) persistent work-table launches in a hipGraph
    groups = [(n, k, bi.ptr, bo.ptr) for (n, k), (bi, bo) in zip(mix, bufs)]
    bt = fa.Batch("float", groups, fa.FFTW_FORWARD, fa.FFTW_ESTIMATE)
    log(bt.info())
    stream = bt.stream()

    def batched():
        bt.execute()

    # the batch runs on its own stream: time it with events on that stream
    bt.execute()
    L.fftw_amd_synchronize()
    t3 = 1e30
    for _ in range(5):
        L.fftw_amd_event_record(e0, stream)
        bt.execute()
        L.fftw_amd_event_record(e1, stream)
        L.fftw_amd_synchronize()
        t3 = min(t3, L.fftw_amd_event_elapsed(e0, e1))
    # accuracy of the batched result: first transform of every size
    errs = {}
    for (n, k), (bi, bo) in zip(mix, bufs):
        x = bi.download(np.complex64, n).astype(np.complex128)
        y = bo.download(np.complex64, n)
        errs[n] = rel_l2(y, np.fft.fft(x))
    byts = 2.0 * pts * 8
    res = {}
    for name, t in (("per_call", t1), ("per_size_batched", t2), ("persistent_graph", t3)):
        res[name] = {"ms": round(t, 3), "transforms_per_s": round(ntr / (t * 1e-3), 1),
                     "GB_s": round(byts / (t * 1e-3) / 1e9, 1)}
        log("%-18s %9.3f ms  %12.1f transforms/s  %8.1f GB/s" % (name, t, ntr / (t * 1e-3),
                                                                  byts / (t * 1e-3) / 1e9))
    rec = {"metric": "transforms/s (uneven-size FP32 c2c set, 1 GPU)",
           "value": res["persistent_graph"]["transforms_per_s"], "unit": "transforms/s",
           "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic seed 3",
           "config": {"workload": "mixed N in {1000,1001,2187,4093,4096,65536,2^20}",
                      "sizes_howmany": mix, "points": pts},
           "speedup_vs_per_call": round(t1 / t3, 2), "speedup_vs_per_size": round(t2 / t3, 2),
           "modes": res, "rel_l2_first_transform": errs}
    print(json.dumps(rec), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only; this engine has no multi-GPU path "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == "mixed":
        run_mixed(args)
        return
    if args.workload == "serve":
        run_serve(args)
        return
    if args.workload == "precision":
        run_precision(args)
        return
    import fftw3_amd as fa

    fa.init()
    L = fa.lib()
    L.fftw_amd_set_async(1)
    n3, n2, H2 = args.n3, args.n2, args.howmany2
    tot3 = n3 ** 3
    nh = n2 // 2 + 1

    # ---- buffers and inputs (device-resident) ----
    log("alloc 3d %d^3 fp64 (%.2f GiB), 2d r2c %dx%d x%d fp32" % (n3, tot3 * 16 / 2**30, n2, n2, H2))
    b3 = fa.DeviceBuffer(tot3 * 16)
    b3.fill_uniform("double", 2 * tot3, 2, 0)
    rin = fa.DeviceBuffer(H2 * n2 * n2 * 4)
    rin.fill_uniform("float", H2 * n2 * n2, 2, 1)
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
    # FFTW_MEASURE overwrote the arrays while timing candidates: refill
    b3.fill_uniform("double", 2 * tot3, 2, 0)
    rin.fill_uniform("float", H2 * n2 * n2, 2, 1)
    log("planned in %.2fs" % (time.time() - t0))
    plans = [("c2c3d_fwd", p3f), ("c2c3d_bwd", p3b), ("r2c2d", pr), ("c2r2d", pc)]
    for name, p in plans:
        for s in p.pass_info():
            log("  %s: %s" % (name, s))

    # ---- correctness of the timed configuration (untimed) ----
    m = 1 << 20
    x3 = counter_uniform(2 * m, 2, 0).view(np.complex128)
    p3f.execute()
    p3b.execute()
    L.fftw_amd_synchronize()
    y3 = b3.download(np.complex128, m)
    rt3 = rel_l2(y3 / tot3, x3)
    b3.fill_uniform("double", 2 * tot3, 2, 0)
    pr.execute()
    pc.execute()
    L.fftw_amd_synchronize()
    xr = rin.download(np.float32, n2 * n2).astype(np.float64)
    yr = rout.download(np.float32, n2 * n2).astype(np.float64)
    rt2 = rel_l2(yr / (n2 * n2), xr)
    log("round trip rel_l2: 3d fp64 %.3e, 2d fp32 %.3e" % (rt3, rt2))

    if args.calib_copy:
        # plain copy of 2^27 complex FP64 (2 GiB read + 2 GiB write): the
        # reference dispatch for scaling FETCH_SIZE / WRITE_SIZE counters
        cp = fa.plan_guru_dft("double", [], [(tot3, 1, 1)], b3.ptr, cmid.ptr if cmid.nbytes >= tot3 * 16
                              else rout.ptr, fa.FFTW_FORWARD)
        tmp = fa.DeviceBuffer(tot3 * 16)
        cp.execute_dft(b3.ptr, tmp.ptr)
        cp.execute_dft(b3.ptr, tmp.ptr)
        L.fftw_amd_synchronize()
        log("calibration copy: 2 dispatches of %d bytes read + %d bytes written" % (tot3 * 16, tot3 * 16))
        del tmp

    # ---- streaming-copy bandwidth of this device (hipMemcpy device to
    # device, 1 GiB read + 1 GiB written), the practical HBM ceiling ----
    copy_gbs = None
    if cmid.nbytes >= (1 << 30) and rout.nbytes >= (1 << 30):
        c0, c1 = L.fftw_amd_event_create(), L.fftw_amd_event_create()
        best = None
        for _ in range(4):
            L.fftw_amd_event_record(c0, None)
            L.fftw_amd_memcpy_dtod(rout.ptr, cmid.ptr, 1 << 30)
            L.fftw_amd_event_record(c1, None)
            L.fftw_amd_synchronize()
            t = L.fftw_amd_event_elapsed(c0, c1)
            best = t if best is None else min(best, t)
        copy_gbs = 2.0 * (1 << 30) / (best * 1e-3) / 1e9
        L.fftw_amd_event_destroy(c0)
        L.fftw_amd_event_destroy(c1)
        log("device copy: %.1f GB/s (1 GiB hipMemcpy, best of 4)" % copy_gbs)

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

    # ---- dominant kernel from per-pass event timing ----
    groups = {}
    sub = {}
    for name, p in plans:
        ip, op = {"c2c3d_fwd": (b3, b3), "c2c3d_bwd": (b3, b3), "r2c2d": (rin, cmid),
                  "c2r2d": (cmid, rout)}[name]
        tot, per = p.time(ip.ptr, op.ptr, iters=3)
        sub[name] = tot
        for info, t, by in zip(p.pass_info(), per, p.pass_bytes()):
            key = info.split(" lines=")[0]
            g = groups.setdefault(key, [0.0, 0.0, 0])
            g[0] += t
            g[1] += by
            g[2] += 1
    key, (tms, tby, cnt) = max(groups.items(), key=lambda kv: kv[1][0])
    achieved = (tby / cnt) / (tms / cnt * 1e-3) / 1e9
    symbol = kernel_symbol(key)
    traffic = committed_traffic(symbol)
    log("dominant kernel: %s  launches/step=%d  mean %.3f ms  %.1f GB/s" % (key, cnt, tms / cnt, achieved))
    for k, (t, by, c) in sorted(groups.items(), key=lambda kv: -kv[1][0]):
        log("  %-90s %8.3f ms %7.1f GB/s x%d" % (k[:90], t, by / (t * 1e-3) / 1e9, c))

    # ---- single-size batched c2c (1-D FP64 N=4096 x 1024, out of place),
    # the configuration of the >= 70 % roofline target; not part of the step
    c1 = single_size_c2c(fa, L)
    dct = r2r_dct2(fa, L)
    log("single-size c2c 4096 x 1024 fp64: %.4f ms  %.1f GB/s  frac %.3f  (%s)"
        % (c1["ms"], c1["GB_s"], c1["frac"], c1["passes"][0]))

    # ---- CPU baseline: oracle double path on the host (bounded sample) ----
    cpu = None
    if not args.no_cpu:
        import oracle  # the CPU baseline is the oracle's double-precision path
        cpu = cpu_baseline(oracle, n3, n2)

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
        "config": {
            "workload": "3-D c2c FP64 %d^3 in-place fwd+bwd + 2-D r2c/c2r FP32 %dx%d x%d" % (n3, n2, n2, H2),
            "parallelism": "single GPU",
            "planner": "FFTW_ESTIMATE" if args.estimate else "FFTW_MEASURE (device autotune)",
            "gflops": round(gflops, 1),
            "bytes_per_step": bytes_step,
            "roundtrip_rel_l2": {"c2c3d_fp64": rt3, "r2c_c2r_2d_fp32": rt2},
            "sub_ms": {k: round(v, 4) for k, v in sub.items()},
            "single_size_c2c_fp64_4096x1024": c1,
            "r2r_redft10_fp32_2048x2048x16": dct,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": symbol,
            "achieved": round(achieved, 1),
            "peak": PEAK_HBM,
            "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM, 4),
            "device_copy_GBps": None if copy_gbs is None else round(copy_gbs, 1),
            "frac_of_copy": None if copy_gbs is None else round(achieved / copy_gbs, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": tby / cnt,
            "mean_launch_ms": round(tms / cnt, 4),
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


def single_size_c2c(fa, L, n=4096, howmany=1024, iters=50):
    """1-D c2c FP64, N=4096, howmany=1024, contiguous, out of place, seed 1
    (FFTW_MEASURE): event-timed mean of `iters` back-to-back executes,
    algorithmic bytes 2*N*howmany*16, and a check of the first transform
    against numpy.fft on the same input."""
    import numpy as np
    tot = n * howmany
    a = fa.DeviceBuffer(tot * 16)
    b = fa.DeviceBuffer(tot * 16)
    a.fill_uniform("double", 2 * tot, 1, 0)
    p = fa.plan_many_dft("double", [n], howmany, a.ptr, None, 1, n, b.ptr, None, 1, n,
                         fa.FFTW_FORWARD, fa.FFTW_MEASURE)
    a.fill_uniform("double", 2 * tot, 1, 0)
    p.execute()
    L.fftw_amd_synchronize()
    x = a.download(np.complex128, n)
    err = rel_l2(b.download(np.complex128, n), np.fft.fft(x))
    e0, e1 = L.fftw_amd_event_create(), L.fftw_amd_event_create()
    for _ in range(3):
        p.execute()
    L.fftw_amd_synchronize()
    L.fftw_amd_event_record(e0, None)
    for _ in range(iters):
        p.execute()
    L.fftw_amd_event_record(e1, None)
    L.fftw_amd_synchronize()
    ms = L.fftw_amd_event_elapsed(e0, e1) / iters
    L.fftw_amd_event_destroy(e0)
    L.fftw_amd_event_destroy(e1)
    gbs = 2.0 * tot * 16 / (ms * 1e-3) / 1e9
    rec = {"ms": round(ms, 5), "GB_s": round(gbs, 1), "frac": round(gbs / PEAK_HBM, 4),
           "rel_l2_first": err, "passes": p.pass_info()}
    p.destroy()
    return rec


def r2r_dct2(fa, L, shape=(2048, 2048), howmany=16, iters=20):
    """REDFT10 (DCT-II) FP32 over every dim of shape x howmany, out of place
    (FFTW_MEASURE): event-timed mean, algorithmic bytes 2*N*howmany*4, and a
    check of the first transform against scipy.fft.dctn (type 2, norm None:
    FFTW's REDFT10 definition) on the same input."""
    import numpy as np
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
    x = a.download(np.float32, N).astype(np.float64).reshape(shape)
    err = rel_l2(b.download(np.float32, N), scipy.fft.dctn(x, type=2).reshape(-1))
    e0, e1 = L.fftw_amd_event_create(), L.fftw_amd_event_create()
    for _ in range(3):
        p.execute()
    L.fftw_amd_synchronize()
    L.fftw_amd_event_record(e0, None)
    for _ in range(iters):
        p.execute()
    L.fftw_amd_event_record(e1, None)
    L.fftw_amd_synchronize()
    ms = L.fftw_amd_event_elapsed(e0, e1) / iters
    L.fftw_amd_event_destroy(e0)
    L.fftw_amd_event_destroy(e1)
    gbs = 2.0 * tot * 4 / (ms * 1e-3) / 1e9
    rec = {"ms": round(ms, 5), "GB_s": round(gbs, 1), "frac": round(gbs / PEAK_HBM, 4),
           "rel_l2_first": err, "passes": p.pass_info()}
    p.destroy()
    return rec


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


def cpu_baseline(oracle, n3, n2):
    """Oracle double path, OpenMP over lines on the host: the 3-D transform
    forward + backward at full size and 4 of the 64 2-D r2c/c2r batches."""
    import numpy as np
    t_all = 0.0
    by = 0.0
    x = oracle.uniform(2 * n3 ** 3, 2, 0).view(np.complex128).reshape(n3, n3, n3)
    t = time.time()
    y = oracle.fftn(x, -1, prec=oracle.D)
    del x
    z = oracle.fftn(y, +1, prec=oracle.D)
    t_all += time.time() - t
    by += 2 * 2.0 * n3 ** 3 * 16
    del y, z
    hs = 4
    xr = oracle.uniform(hs * n2 * n2, 2, 1).reshape(hs, n2, n2)
    t = time.time()
    X = oracle.rfftn(xr, axes_batch=1, prec=oracle.D)
    yr = oracle.irfftn(X, n2, axes_batch=1, prec=oracle.D)
    t_all += time.time() - t
    by += 2 * hs * (n2 * n2 * 4.0 + n2 * (n2 // 2 + 1) * 8.0)
    del X, yr, xr
    return {"value": round(by / t_all / 1e9, 3), "unit": "GB/s", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": "oracle double-precision path: 3-D %d^3 fwd+bwd (full) + 2-D r2c+c2r "
                      "%dx%d for 4 of 64 batches; %.1f s" % (n3, n2, n2, t_all)}


if __name__ == "__main__":
    main()
