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
    ap.add_argument("--no-batch", action="store_true", help="skip the batched uneven-slices leg")
    ap.add_argument("--no-dda", action="store_true", help="skip the line-traversal (Bresenham) leg")
    ap.add_argument("--batch-slices", type=int, default=256, help="slices in the batched leg")
    ap.add_argument("--batch-iters", type=int, default=20, help="OSEM iterations in the batched leg")
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


def reference_case(dev, reps=50):
    """The reference's own CPU-runnable case: 256x256 Shepp-Logan (10 ellipses), 180 angles
    in [0, pi), 256 bins over [-1, 1].  Radon, backprojection and FBP on the GPU (device-
    resident, HIP events on the launching stream, mean of `reps` after a warm-up); the
    reference C's time for the same case is in cpu_baseline."""
    import numpy as np
    import torch
    import pyraft_amd as pr
    from pyraft_amd.device import Plan
    n, R, A = 256, 256, 180
    f = np.asarray(pr.shepplogan((n, n)), dtype=np.float32)
    p = Plan(n, R, A)
    x = torch.from_numpy(f).to(dev)
    s = p.radon(x)

    def ev(fn):
        fn()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.synchronize()
        return round(a.elapsed_time(b) / reps * 1e3, 2)

    return {"workload": "256x256 Shepp-Logan, 180 angles in [0,pi), 256 bins over [-1,1] (the reference's "
                        "CPU-runnable case); device-resident, mean of %d" % reps,
            "radon_us": ev(lambda: p.radon(x)), "backprojection_us": ev(lambda: p.backprojection(s)),
            "fbp_us": ev(lambda: p.fbp(s))}


def _cpu_reference_case(O, ref, nthreads):
    """The reference C (or the oracle) on the 256^2 Shepp-Logan case: median of 3."""
    import numpy as np
    import pyraft_amd as pr
    n, R, A = 256, 256, 180
    img = np.asarray(pr.shepplogan((n, n)), dtype=np.float32).astype(np.float64)
    radon = (lambda: O.ref_radon(img, R, A, nthreads=nthreads)) if ref else (
        lambda: O.radon(img, R, A, nthreads=nthreads))
    bp = (lambda y: O.ref_backprojection(y, n, nthreads=nthreads)) if ref else (
        lambda y: O.backprojection(y, n, nthreads=nthreads))
    sino = radon()
    tr, tb = [], []
    for _ in range(3):
        t0 = time.perf_counter()
        sino = radon()
        t1 = time.perf_counter()
        bp(sino)
        tb.append(time.perf_counter() - t1)
        tr.append(t1 - t0)
    return {"workload": "256x256 Shepp-Logan, 180 angles, 256 bins", "radon_us": round(statistics.median(tr) * 1e6, 1),
            "backprojection_us": round(statistics.median(tb) * 1e6, 1)}


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
        "reference_case": _cpu_reference_case(O, kind == "reference", nthreads),
    }


def batched_leg(dev, nslices, niter, nsubsets=8, seed=4):
    """Uneven slices: n from {256,...,2048}, A ~ U[180,1800], R = n, 32-ellipse phantoms.
    Radon + backprojection and OSEM (nsubsets x niter, noiseless data s = R f), each as
    (a) one batched persistent launch per kernel captured in a HIP graph, after the
    autotuner picked the backprojection variant per image size, and (b) plain
    per-slice launches of the same operators with the default variant."""
    import numpy as np
    import torch
    from pyraft_amd.batch import Batch
    from pyraft_amd.device import Plan, phantom
    from pyraft_amd.phantom import random_ellipses

    rng = np.random.default_rng(seed)
    sizes = [int(v) for v in rng.choice([256, 512, 768, 1024, 1536, 2048], nslices)]
    angles = [int(v) for v in rng.integers(180, 1801, nslices)]
    upd = float(sum(n * n * a for n, a in zip(sizes, angles)))
    print("bench: batch of %d slices, %.3g updates per pass" % (nslices, upd), file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    b = Batch(sizes, sizes, angles)
    plans = {}
    for n, a in zip(sizes, angles):
        if (n, a) not in plans:
            plans[(n, a)] = Plan(n, n, a)
    t_plan = time.perf_counter() - t0
    images = b.empty_images()
    for i, n in enumerate(sizes):
        b.image(images, i).copy_(phantom(random_ellipses(32, 1000 + i), n))
    t0 = time.perf_counter()
    variants = b.tune(reps=1)
    torch.cuda.synchronize()
    t_tune = time.perf_counter() - t0
    tuned = {}
    for n, v in zip(sizes, variants):
        tuned[str(n)] = v

    # ---- Radon + backprojection pass
    sinos = b.empty_sinos()
    bps = b.empty_images()
    b.radon(images, out=sinos, graph=True)          # capture
    b.backprojection(sinos, out=bps, graph=True)
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        b.radon(images, out=sinos, graph=True)
        b.backprojection(sinos, out=bps, graph=True)
    torch.cuda.synchronize()
    t_batched = (time.perf_counter() - t0) / reps
    print("bench: batched Radon+BP %.1f ms" % (t_batched * 1e3), file=sys.stderr, flush=True)
    views = [b.image(images, i) for i in range(nslices)]
    plain_sinos = [None] * nslices
    plain_bps = [None] * nslices

    def plain_pass():
        for i, (n, a) in enumerate(zip(sizes, angles)):
            p = plans[(n, a)]
            plain_sinos[i] = p.radon(views[i])
            plain_bps[i] = p.backprojection(plain_sinos[i])

    plain_pass()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plain_pass()
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / reps
    print("bench: plain Radon+BP %.1f ms" % (t_plain * 1e3), file=sys.stderr, flush=True)

    def worst(a_list, b_list):
        w = 0.0
        for x, y in zip(a_list, b_list):
            d = float(torch.linalg.vector_norm((x - y).double()))
            r = float(torch.linalg.vector_norm(y.double()))
            w = max(w, d / (r if r > 0 else 1.0))
        return w

    diff_sino = worst([b.sino(sinos, i) for i in range(nslices)], plain_sinos)
    diff_bp = worst([b.image(bps, i) for i in range(nslices)], plain_bps)

    # ---- OSEM
    x = torch.ones(b.img_total, dtype=torch.float32, device=dev)
    b.em(sinos, 1, nsubsets, x=x)                   # weights, tables, kernels
    x.fill_(1.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.em(sinos, niter, nsubsets, x=x, graph=True)    # capture + one graph launch
    torch.cuda.synchronize()
    t_em_b = time.perf_counter() - t0
    print("bench: batched OSEM %.2f s" % t_em_b, file=sys.stderr, flush=True)
    plain_x = [None] * nslices
    plans[(sizes[0], angles[0])].em(plain_sinos[0], 1, nsubsets)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i, (n, a) in enumerate(zip(sizes, angles)):
        plain_x[i] = plans[(n, a)].em(plain_sinos[i], niter, nsubsets)
    torch.cuda.synchronize()
    t_em_p = time.perf_counter() - t0
    print("bench: plain OSEM %.2f s" % t_em_p, file=sys.stderr, flush=True)
    diff_em = worst([b.image(x, i) for i in range(nslices)], plain_x)
    return {
        "workload": "%d slices, n in {256..2048} (seed %d), A ~ U[180,1800], R = n, 32-ellipse phantoms; "
                    "Radon+backprojection pass and OSEM %d subsets x %d iterations on s = R f"
                    % (nslices, seed, nsubsets, niter),
        "updates_per_pass": upd,
        "radon_bp_ms_batched": round(t_batched * 1e3, 3),
        "radon_bp_ms_plain": round(t_plain * 1e3, 3),
        "radon_bp_gups_batched": round(2 * upd / t_batched / 1e9, 2),
        "radon_bp_gups_plain": round(2 * upd / t_plain / 1e9, 2),
        "radon_bp_speedup": round(t_plain / t_batched, 3),
        "em_iters_per_s_batched": round(niter / t_em_b, 3),
        "em_iters_per_s_plain": round(niter / t_em_p, 3),
        "em_speedup": round(t_em_p / t_em_b, 3),
        "em_gups_batched": round(2 * upd * niter / t_em_b / 1e9, 2),
        "tuned_bp_variant_by_n": tuned,
        "tune_s": round(t_tune, 3),
        "plan_s": round(t_plan, 3),
        "max_rel_l2_batched_vs_plain": {"radon": diff_sino, "backprojection": diff_bp, "osem": diff_em},
    }


def dda_leg(dev, img, n, R, A, reps=3):
    """pyraft's default projector pair (line traversal) on the headline image and angles:
    event-timed Radon, transpose and normalised backprojection (plan and weights built
    beforehand)."""
    print("bench: line-traversal pair", file=sys.stderr, flush=True)
    import torch
    from pyraft_amd.device import DdaPlan
    plan = DdaPlan((n, n), R, A)
    y = torch.rand((R, A), device=dev, generator=torch.Generator(dev).manual_seed(9))
    out = {"workload": "%d^2 image, %d rays x %d angles, theta in [0, pi] (A - 1 steps), FP64 traversal, "
                       "FP32 taps" % (n, R, A)}
    info = plan.info()
    out["traced_angles"] = info["traced_angles"]
    ups = float(n) * n * A
    for name, fn in (("radon", lambda: plan.radon(img)), ("transpose", lambda: plan.radon_transpose(y)),
                     ("backprojection", lambda: plan.backprojection(y))):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out[name + "_ms"] = round(ms, 3)
        out[name + "_gups"] = round(ups / ms / 1e6, 1)
    return out


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
    bp_variant, bp_variant_us = plan.tune(reps=3)   # autotuned backprojection tile / LDS stage
    from pyraft_amd.device import bp_variant_shape
    bp_rows, bp_stage = bp_variant_shape(bp_variant)
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

    # exact transpose of the Radon (the autograd backward) on the same sinogram
    for _ in range(2):
        plan.radon_transpose(sino)
    e0.record(stream)
    for _ in range(3):
        plan.radon_transpose(sino)
    e1.record(stream)
    torch.cuda.synchronize()
    transpose_ms = e0.elapsed_time(e1) / 3

    dominant = ("k_backprojection<0, %d, %d>" % (bp_rows, bp_stage)) if t_bp >= t_radon else "k_radon<1>"
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
        "exact_transpose_ms": round(transpose_ms, 4),
        "bp_variant": {"chosen": bp_variant, "rows_per_thread": bp_rows, "stage_pairs": bp_stage,
                       "tuned_us": {str(k): round(v, 1) for k, v in bp_variant_us.items()}},
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
        # the same OSEM through the batch path: one persistent launch per kernel,
        # the whole call captured once into a HIP graph and replayed
        from pyraft_amd.batch import Batch
        eb = Batch([ne], [ne], [Ae])
        eb.tune(reps=2)
        sb = eb.pack_sinos([s])
        xb = torch.ones(eb.img_total, dtype=torch.float32, device=dev)
        eb.em(sb, iters, S, x=xb, graph=True)          # capture + first replay
        xb.fill_(1.0)
        torch.cuda.synchronize()
        td = time.perf_counter()
        _, llb = eb.em(sb, iters, S, x=xb, return_ll=True, graph=True)
        torch.cuda.synchronize()
        te = time.perf_counter()
        it_b = iters / (te - td)
        xd = eb.image(xb, 0).double()
        em_diff = float(torch.linalg.vector_norm(xd - x.double()) / torch.linalg.vector_norm(x.double()))
        result["em"] = {"workload": "OSEM 1024x1024, 720 angles, 1024 bins, 16 subsets, %d iterations, "
                                    "Poisson data (mean count 1e3)" % iters,
                        "iters_per_s_plain": round(it_s, 2),
                        "iters_per_s_batched_graph": round(it_b, 2),
                        "speedup": round(it_b / it_s, 3),
                        "gups": round(2.0 * ne * ne * Ae * it_s / 1e9, 2),
                        "gups_batched_graph": round(2.0 * ne * ne * Ae * it_b / 1e9, 2),
                        "rel_l2_batched_vs_plain": em_diff,
                        "ll_first_last": [round(ll[0], 3), round(ll[-1], 3)],
                        "ll_first_last_batched": [round(llb[0][0], 3), round(llb[-1][0], 3)]}

    # the drop-in host path (pyraft.radon / backprojection: FP64 numpy in and out,
    # PCIe transfers and FP64<->FP32 conversion included), same workload
    print("bench: host-buffer drop-in path", file=sys.stderr, flush=True)
    import pyraft_amd as pr
    # a pyraft image (default corners [-1,1]^2): a bare ndarray would be read as an
    # ellipse description, as pyraft's radon does (pyraft/sinogram/sino.py:175-212)
    host_img = pr.image(img.double().cpu().numpy())
    pr.radon(host_img, (R, A), method=pr.radon_method.SLANTSTACK)  # plan cache, warm-up
    t0 = time.perf_counter()
    host_sino = pr.radon(host_img, (R, A), method=pr.radon_method.SLANTSTACK)
    t1 = time.perf_counter()
    pr.backprojection(host_sino, pr.backprojection_plan(host_sino.shape, (n, n), method=pr.radon_method.SLANTSTACK))
    t2 = time.perf_counter()
    result["host_path"] = {"note": "pyraft-compatible radon/backprojection on FP64 host arrays (H2D/D2H, "
                                   "FP64<->FP32 conversion included); value is device-resident",
                           "radon_ms": round((t1 - t0) * 1e3, 3), "backprojection_ms": round((t2 - t1) * 1e3, 3),
                           "gups": round(2 * upd / (t2 - t0) / 1e9, 2)}

    # differentiable projector: forward + backward (exact transpose) on 64 x 512^2, 720 angles
    print("bench: autograd 64 x 512^2 x 720", file=sys.stderr, flush=True)
    from pyraft_amd.autograd import radon as ad_radon
    aplan = Plan(512, 512, 720)
    xa = torch.rand((64, 512, 512), device=dev, requires_grad=True)
    ya = torch.rand((64, 512, 720), device=dev)
    for _ in range(2):
        ((ad_radon(xa, aplan) * ya).sum()).backward()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        ((ad_radon(xa, aplan) * ya).sum()).backward()
    torch.cuda.synchronize()
    ta = (time.perf_counter() - t0) / 3
    aupd = 64.0 * 512 * 512 * 720
    result["autograd"] = {"workload": "64 x 512^2, 720 angles: Radon forward + exact-transpose backward "
                                      "(+ torch elementwise/sum)",
                          "ms_per_step": round(ta * 1e3, 3), "gups": round(2 * aupd / ta / 1e9, 2)}

    if not args.no_dda:
        result["bresenham"] = dda_leg(dev, img, n, R, A)

    if not args.no_batch:
        result["batched"] = batched_leg(dev, args.batch_slices, args.batch_iters)

    print("bench: reference case (256^2 Shepp-Logan, 180 angles)", file=sys.stderr, flush=True)
    result["reference_case"] = reference_case(dev)

    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(n, A)

    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
