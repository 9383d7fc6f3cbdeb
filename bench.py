#!/usr/bin/env python3
"""Benchmark: slant-stack Radon + backprojection GUPS at 2048^2 x 1500 angles (FP32).

One step = one slant-stack Radon of a 2048x2048 image into a 1500-angle x 2048-bin
sinogram plus one backprojection of that sinogram, both on the GPU with inputs
already resident in HBM.  GUPS = (N^2 * A updates per operator, both operators)
/ time.  Extra fields: per-operator GUPS, FBP time, EM/OSEM iterations per second
on the 1024^2 / 720-angle / 16-subset workload, the roofline of the dominant
kernel and the reference's CPU path timed on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.
"""
import argparse
import json
import math
import os
import platform
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Radon/backproj GUPS at 2048²×1500 angles; EM iters/s batched vs plain, FP32"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
BYTES_PER_UPDATE = 8.0         # two FP32 interpolation taps per linear-interpolated accumulate


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=2048)
    ap.add_argument("--angles", type=int, default=1500)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-em", action="store_true", help="skip the EM leg")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="run one 64 MiB device-to-device copy first (reference for HBM byte counters)")
    return ap.parse_args()


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` at this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or None."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "hbm_bytes_*2048x1500*.md"))):
        m = re.search(r"```json\n(.*?)\n```", open(path).read(), flags=re.S)
        if not m:
            continue
        table = json.loads(m.group(1)).get("bytes_per_launch", {})
        for key, val in table.items():
            if key.split("/")[0] == kernel:
                return float(val), os.path.relpath(path, ROOT)
    return None, None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(n, angles):
    """Reference C (compiled from the reference sources into oracle/_ref) on a
    bounded sample of the same workload: the full 2048^2 image with a uniformly
    spread subset of the angles, Radon then backprojection, nthreads = cpu_count()
    as pyraft sets it.  Falls back to the oracle restatement when _ref is absent."""
    import numpy as np
    from oracle import oracle as O
    from pyraft_amd.phantom import random_ellipses
    nthreads = os.cpu_count() or 1
    sub = 24
    img = O.phantom_fromdesc(random_ellipses(32, 2), n).astype(np.float32).astype(np.float64)
    kind = "reference" if O.ref_available() else "port"
    if kind == "reference":
        radon = lambda: O.ref_radon(img, n, sub, nthreads=nthreads)
        bp = lambda s: O.ref_backprojection(s, n, nthreads=nthreads)
    else:
        radon = lambda: O.radon(img, n, sub, nthreads=nthreads)
        bp = lambda s: O.backprojection(s, n, nthreads=nthreads)
    times = []
    sino = radon()  # warm-up
    for _ in range(3):
        t0 = time.perf_counter()
        sino = radon()
        t1 = time.perf_counter()
        bp(sino)
        t2 = time.perf_counter()
        times.append((t1 - t0, t2 - t1))
    tr = statistics.median(t[0] for t in times)
    tb = statistics.median(t[1] for t in times)
    upd = float(n) * n * sub
    return {
        "value": round(2 * upd / (tr + tb) / 1e9, 4),
        "unit": "GUPS",
        "cores": nthreads,
        "kind": kind,
        "sample": "%dx%d image, %d of the %d angles (uniform over [0,pi)), %d bins; Radon+backprojection, "
                  "median of 3 after 1 warm-up; FP64 reference C, nthreads=cpu_count()" % (n, n, sub, angles, n),
        "radon_gups": round(upd / tr / 1e9, 4),
        "backprojection_gups": round(upd / tb / 1e9, 4),
        "cpu_model": cpu_model(),
    }


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build measures one GPU only (got --gpus %d, WORLD_SIZE %d); "
              "run with --gpus 1" % (args.gpus, world), file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from pyraft_amd.device import Plan, phantom
    from pyraft_amd.phantom import random_ellipses

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    n, A = args.n, args.angles
    R = n
    upd = float(n) * n * A

    if args.calibration_copy:
        src = torch.ones(16 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    print("bench: plan n=%d R=%d A=%d" % (n, R, A), file=sys.stderr, flush=True)
    plan = Plan(n, R, A)
    img = phantom(random_ellipses(32, 2), n)
    sinoT = torch.empty((A, R), dtype=torch.float32, device=dev)
    bp = torch.empty((n, n), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        plan.radon(img, angle_major=True, out=sinoT)
        if ev is not None:
            ev[1].record(stream)
        out = plan.backprojection(sinoT, angle_major=True)
        if ev is not None:
            ev[2].record(stream)
        return out

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    print("bench: timing %d steps" % args.steps, file=sys.stderr, flush=True)
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        bp = step(events[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    t_radon = statistics.mean(e[0].elapsed_time(e[1]) for e in events) * 1e-3
    t_bp = statistics.mean(e[1].elapsed_time(e[2]) for e in events) * 1e-3
    ms_per_step = wall / args.steps * 1e3
    value = 2 * upd / (wall / args.steps) / 1e9
    if not (torch.isfinite(sinoT).all() and torch.isfinite(bp).all()):
        print("bench: non-finite output", file=sys.stderr)
        sys.exit(1)

    # FBP (filter + backprojection) on the same sinogram, reference layout
    sino = sinoT.t().contiguous()
    fbp_ms = None
    for _ in range(2):
        plan.fbp(sino)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(3):
        plan.fbp(sino)
    e1.record(stream)
    torch.cuda.synchronize()
    fbp_ms = e0.elapsed_time(e1) / 3

    dominant = "k_backprojection<0>" if t_bp >= t_radon else "k_radon<1>"
    t_dom = max(t_bp, t_radon)
    achieved = BYTES_PER_UPDATE * upd / t_dom / 1e9
    traffic, traffic_src = measured_traffic(dominant)
    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: 32 seeded random ellipses (seed 2), rasterised on the GPU",
        "config": {"workload": "slant-stack Radon + backprojection, %dx%d image, %d angles in [0,pi), %d bins"
                               % (n, n, A, R), "n": n, "angles": A, "rays": R},
        "radon_gups": round(upd / t_radon / 1e9, 2),
        "backprojection_gups": round(upd / t_bp / 1e9, 2),
        "radon_ms": round(t_radon * 1e3, 4),
        "backprojection_ms": round(t_bp * 1e3, 4),
        "fbp_ms": round(fbp_ms, 4),
        "roofline": {"bound": "hbm", "kernel": dominant, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "model": "8 B (two FP32 taps) per update x N^2*A updates per launch / mean launch time; "
                              "traffic = measured HBM bytes per launch (copy-calibrated FETCH_SIZE + WRITE_SIZE)"},
    }

    if not args.no_em:
        print("bench: EM/OSEM 1024^2, 720 angles, 16 subsets", file=sys.stderr, flush=True)
        ne, Ae, S, iters = 1024, 720, 16, 50
        eplan = Plan(ne, ne, Ae)
        f = phantom(random_ellipses(32, 3), ne)
        p = eplan.radon(f)
        p = p * (1e3 / p.mean())
        rng = np.random.default_rng(30)
        s = torch.from_numpy(rng.poisson(p.cpu().numpy().astype(np.float64)).astype(np.float32)).to(dev)
        eplan.em(s, 1, S)  # warm-up (workspace, kernels)
        torch.cuda.synchronize()
        ta = time.perf_counter()
        eplan.em(s, 0, S)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        x, ll = eplan.em(s, iters, S, return_ll=True)
        torch.cuda.synchronize()
        tc = time.perf_counter()
        it_s = iters / max(1e-9, (tc - tb) - (tb - ta))
        result["em"] = {"workload": "OSEM 1024x1024, 720 angles, 1024 bins, 16 subsets, %d iterations, "
                                    "Poisson data (mean count 1e3)" % iters,
                        "iters_per_s_plain": round(it_s, 2),
                        "gups": round(2.0 * ne * ne * Ae * it_s / 1e9, 2),
                        "ll_first_last": [round(ll[0], 3), round(ll[-1], 3)]}

    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(n, A)

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
