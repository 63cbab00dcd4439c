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
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only; this build has no multi-GPU path "
              "(requested --gpus %d)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import fftw3_amd as fa
    import oracle  # checker only: round-trip references and cpu_baseline

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
    x3 = oracle.uniform(2 * m, 2, 0).view(np.complex128)
    p3f.execute()
    p3b.execute()
    L.fftw_amd_synchronize()
    y3 = b3.download(np.complex128, m)
    rt3 = oracle.rel_l2(y3 / tot3, x3)
    b3.fill_uniform("double", 2 * tot3, 2, 0)
    pr.execute()
    pc.execute()
    L.fftw_amd_synchronize()
    xr = rin.download(np.float32, n2 * n2).astype(np.float64)
    yr = rout.download(np.float32, n2 * n2).astype(np.float64)
    rt2 = oracle.rel_l2(yr / (n2 * n2), xr)
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

    # ---- CPU baseline: oracle double path on the host (bounded sample) ----
    cpu = None
    if not args.no_cpu:
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
        },
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
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(rec), flush=True)


def kernel_symbol(info):
    """Demangled kernel name of a Stockham pass from its pass description
    ("stockham f64 L=512 rad=8x8x8 col plain tpl=64 nl=8")."""
    f = info.split()
    if f[0] != "stockham":
        return info
    kv = dict(x.split("=", 1) for x in f if "=" in x)
    rads = kv["rad"].split("x")
    io = {"plain": 0, "tout": 1, "r2c": 2, "c2r": 3, "bluestein": 4}[f[5]]
    return "void fk::stockham<%s, %s, %d, %d, %s, fk::RadList<%s>, %s>(fk_args)" % (
        "double" if f[1] == "f64" else "float", kv["nl"], 0 if f[4] == "row" else 1, io, rads[0],
        ", ".join(rads), kv["tpl"])


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
