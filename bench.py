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
LDS_PEAK_GBS = 78643.2         # LDS reads: 256 CUs x 128 B per clock x 2.4 GHz
VALU_PEAK_TFLOPS = 157.3       # MI355X FP32 vector (published; counts packed FP32)
CUS, CLOCK_HZ = 256, 2.4e9     # MI355X compute units and peak engine clock
BYTES_PER_UPDATE = 8.0         # two FP32 interpolation taps per linear-interpolated accumulate
FLOPS_PER_UPDATE = 6.0         # coordinate FMA (2), lerp (sub + FMA, 3), accumulate (1)


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
    ap.add_argument("--no-rows", action="store_true",
                    help="skip the mesh Radon, transmission EM and binary file legs")
    ap.add_argument("--batch-slices", type=int, default=256, help="slices in the batched leg")
    ap.add_argument("--batch-iters", type=int, default=20, help="OSEM iterations in the batched leg")
    ap.add_argument("--calibration-copy", action="store_true",
                    help="run one 64 MiB device-to-device copy first (reference for HBM byte counters)")
    return ap.parse_args()


def roofline_of(kernel, updates, seconds, traffic, traffic_src, counters=None):
    """Roofline of the dominant kernel against the on-chip unit its algorithm loads.

    Both kernels read their interpolation taps from LDS (8 B, one (tap, difference)
    record or staged pair, per update) and do ~6 FP32 operations per update; HBM traffic
    is a small fraction of the tap bytes because windows are staged once per block.
    `frac` is the algorithmic demand on the unit over the unit's peak, <= 1 by
    construction:
      lds (the Radon):          8 B x updates / time  over  LDS_PEAK_GBS
      valu (the backprojection): 6 flops x updates / time  over  VALU_PEAK_TFLOPS
    VALU_PEAK_TFLOPS is the published FP32 figure, which counts packed FP32; these kernels
    are built without packed FP32, so `valu.frac_nonpacked` (over half of it) is the
    ceiling they can reach.  Beside it: the counters of the committed rocprofv3 summary
    for this kernel (VALU / SALU / LDS wave-instructions per update and per CU-clock,
    LDS bank-conflict cycles over LDS-busy cycles), the measured HBM bytes (FETCH_SIZE +
    WRITE_SIZE, copy-calibrated) over the HBM peak, and the uncached-gather model (8 B
    per update from HBM), which exceeds 1 because the taps come from LDS."""
    lds_gbs = BYTES_PER_UPDATE * updates / seconds / 1e9
    tflops = FLOPS_PER_UPDATE * updates / seconds / 1e12
    if kernel.startswith("k_radon"):
        main = {"bound": "lds", "achieved": round(lds_gbs, 1), "peak": LDS_PEAK_GBS, "unit": "GB/s",
                "frac": round(lds_gbs / LDS_PEAK_GBS, 4)}
    else:
        main = {"bound": "valu", "achieved": round(tflops, 2), "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tflops / VALU_PEAK_TFLOPS, 4)}
    main.update({
        "kernel": kernel,
        "traffic": traffic,
        "traffic_source": traffic_src,
        "model": "updates = N^2 * A per launch; LDS: 8 B (the tap pair) per update; VALU: 6 FP32 flops per update "
                 "(coordinate FMA, lerp, accumulate); time = mean event-timed launch of the operator",
        "lds": {"achieved_gbs": round(lds_gbs, 1), "peak_gbs": LDS_PEAK_GBS, "frac": round(lds_gbs / LDS_PEAK_GBS, 4)},
        "valu": {"achieved_tflops": round(tflops, 2), "peak_tflops": VALU_PEAK_TFLOPS,
                 "frac": round(tflops / VALU_PEAK_TFLOPS, 4),
                 "frac_nonpacked": round(tflops / (VALU_PEAK_TFLOPS / 2), 4)},
        "hbm_measured": None if traffic is None else {
            "achieved_gbs": round(traffic / seconds / 1e9, 1), "peak_gbs": HBM_PEAK_GBS,
            "frac": round(traffic / seconds / 1e9 / HBM_PEAK_GBS, 4)},
        "uncached_gather_model": {"achieved_gbs": round(lds_gbs, 1), "peak_gbs": HBM_PEAK_GBS,
                                  "ratio": round(lds_gbs / HBM_PEAK_GBS, 3),
                                  "note": "8 B per update as if every tap came from HBM; exceeds 1 because the "
                                          "taps are served from LDS"},
    })
    if counters:
        cu_clk = seconds * CUS * CLOCK_HZ
        c = {"source": counters["source"], "launch": counters.get("launch")}
        for name, key in (("valu", "SQ_INSTS_VALU"), ("salu", "SQ_INSTS_SALU"), ("lds", "SQ_INSTS_LDS")):
            if key in counters:
                c[name + "_lane_ops_per_update"] = round(counters[key] * 64 / updates, 3)
                c[name + "_insts_per_cu_clock"] = round(counters[key] / cu_clk, 3)
        if "SQ_LDS_BANK_CONFLICT" in counters and counters.get("SQ_LDS_IDX_ACTIVE"):
            c["lds_bank_conflict_share"] = round(counters["SQ_LDS_BANK_CONFLICT"] / counters["SQ_LDS_IDX_ACTIVE"], 4)
        main["counters"] = c
    return main


def _profile_name_matches(key, kernel):
    """A profile key ("name/work-items") names `kernel`; a kernel given without its last
    template arguments ("k_radon<1>") matches every instantiation that extends it
    ("k_radon<1, 1>": the Radon's block shape is picked per launch)."""
    name = key.split("/")[0]
    return name == kernel or (kernel.endswith(">") and name.startswith(kernel[:-1] + ", "))


def _profile_json(pattern, field):
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", pattern))):
        m = re.search(r"```json\n(.*?)\n```", open(path).read(), flags=re.S)
        if m:
            yield os.path.relpath(path, ROOT), json.loads(m.group(1)).get(field, {})


def measured_counters(kernel):
    """SQ counters per launch of `kernel` at this workload from the committed rocprofv3
    summary (profiles/counters_*2048x1500*.md), or None: of the matching launch shapes the
    one with the most VALU instructions (the full-size launch)."""
    for src, table in _profile_json("counters_*2048x1500*.md", "per_launch"):
        hits = [(key, val) for key, val in table.items() if _profile_name_matches(key, kernel)]
        if hits:
            key, val = max(hits, key=lambda kv: kv[1].get("SQ_INSTS_VALU", 0))
            out = dict(val)
            out["source"] = src
            out["launch"] = key
            return out
    return None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` at this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-calibrated), or None: of the matching
    launch shapes the one moving the most bytes."""
    for src, table in _profile_json("hbm_bytes_*2048x1500*.md", "bytes_per_launch"):
        hits = [(key, float(val)) for key, val in table.items() if _profile_name_matches(key, kernel)]
        if hits:
            return max(v for _, v in hits), src
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


def _event_ms(fn, reps, stream):
    """Mean milliseconds per call of `reps` back-to-back calls, HIP events on `stream`."""
    import torch
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(reps):
        fn()
    b.record(stream)
    b.synchronize()
    return a.elapsed_time(b) / reps


def _graph_ms(fn, reps):
    """The same call captured once into a HIP graph and replayed `reps` times."""
    import torch
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(side):
        fn()
        torch.cuda.synchronize()
        with torch.cuda.graph(g, stream=side):
            fn()
        g.replay()
        torch.cuda.synchronize()
        ms = _event_ms(g.replay, reps, side)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    return ms


def reference_case(dev, reps=50):
    """The reference's own CPU-runnable case: 256x256 Shepp-Logan (10 ellipses), 180 angles
    in [0, pi), 256 bins over [-1, 1].  Radon, backprojection and FBP on the GPU, device-
    resident: `*_us` is the mean of `reps` calls enqueued from Python (host launch cost
    included), `*_graph_us` the mean of `reps` replays of the call captured in a HIP graph
    (a graph launch per call costs more on this runtime than the call's own kernel
    launches, so it is reported beside the eager time, not as it).  The reference C's
    time for the same case is in cpu_baseline."""
    import numpy as np
    import torch
    import pyraft_amd as pr
    from pyraft_amd.device import Plan
    n, R, A = 256, 256, 180
    f = np.asarray(pr.shepplogan((n, n)), dtype=np.float32)
    p = Plan(n, R, A)
    x = torch.from_numpy(f).to(dev)
    s = p.radon(x)
    ops = {"radon": lambda: p.radon(x), "backprojection": lambda: p.backprojection(s), "fbp": lambda: p.fbp(s)}
    out = {"workload": "256x256 Shepp-Logan, 180 angles in [0,pi), 256 bins over [-1,1] (the reference's "
                       "CPU-runnable case); device-resident, mean of %d calls" % reps}
    stream = torch.cuda.current_stream()
    for name, fn in ops.items():
        fn()
        torch.cuda.synchronize()
        out[name + "_us"] = round(_event_ms(fn, reps, stream) * 1e3, 2)
    for name, fn in ops.items():
        out[name + "_graph_us"] = round(_graph_ms(fn, reps) * 1e3, 2)
    return out


def _cpu_ops(O, ref, nthreads):
    """Radon and backprojection of the reference C (oracle/_ref) or, without it, of the
    oracle's restatement, on an angle block [j0, j1) of an A-angle grid over [0, pi)
    (corners of the block: theta_j0 and theta_j0 + (j1 - j0) * pi / A, so the block's
    angles are the full grid's)."""
    import numpy as np

    def corners(A, j0, j1):
        dth = math.pi / A
        return (j0 * dth, 1.0), (j0 * dth + (j1 - j0) * dth, -1.0)

    if ref:
        def radon(img, R, A, j0=0, j1=None):
            j1 = A if j1 is None else j1
            tl, br = corners(A, j0, j1)
            return O.ref_radon(img, R, j1 - j0, sino_tl=tl, sino_br=br, nthreads=nthreads)

        def bp(sino_block, n, A, j0=0):
            tl, br = corners(A, j0, j0 + sino_block.shape[1])
            return O.ref_backprojection(sino_block, n, sino_tl=tl, sino_br=br, nthreads=nthreads)
    else:
        def radon(img, R, A, j0=0, j1=None):
            j1 = A if j1 is None else j1
            return O.radon(img, R, A, j0=j0, j1=j1, nthreads=nthreads)[:, j0:j1]

        def bp(sino_block, n, A, j0=0):
            full = np.zeros((sino_block.shape[0], A))
            full[:, j0:j0 + sino_block.shape[1]] = sino_block
            return O.backprojection(full, n, j0=j0, j1=j0 + sino_block.shape[1], nthreads=nthreads)
    return radon, bp


def cpu_baseline(n, angles):
    """The reference's CPU path on this host's cores: its own C (slant-stack Radon and
    backprojection, FP64, compiled from the reference sources into oracle/_ref) or, where
    that library was not built, the oracle's restatement of it (kind "port").  Threads:
    min(16, cpu_count()), the GPU's share of the host (pyraft itself passes cpu_count()).

    * 2048^2 x 1500: Radon and backprojection in full, one run each (the timed value);
    * the reference's 256^2 Shepp-Logan case: Radon, backprojection and FBP (numpy ramp
      filter, pyraft/filters/filters.py:36-68, then backprojection), median of 3;
    * one OSEM iteration at 1024^2 / 720 angles / 16 subsets composed from the C operators
      with the raft_em_gpu recurrence (raft/raft_em_gpu.cu:41-69): per subset a Radon of
      its 45 angles, the ratio, a backprojection and the update (weights precomputed);
    * the line-traversal Radon and transpose (the reference's iterator) on 24 angles, and
      the reference's binary image write and read of the 2048^2 image;
    * the mesh Radon (the reference's withmesh) on 24 of the mesh leg's angles, and one
      transmission-EM iteration at 1024^2 / 720 angles composed from the C operators with
      the raft_tr_em_gpu recurrence (raft/raft_transmission_em.cu:41-59);
    * the 256-slice uneven batch from a stratified sample: per image side the slices with
      the fewest and the most angles, each projected and backprojected on 24 of its
      angles; the per-update rate of each side extrapolates to the whole pass and to
      OSEM 8 subsets x 20 iterations (each iteration one Radon and one backprojection
      pass)."""
    import numpy as np
    from oracle import oracle as O
    from pyraft_amd.phantom import random_ellipses
    nthreads = min(16, os.cpu_count() or 1)
    ref = O.ref_available()
    kind = "reference" if ref else "port"
    radon, bp = _cpu_ops(O, ref, nthreads)
    out = {"unit": "GUPS", "cores": nthreads, "kind": kind, "cpu_model": cpu_model()}
    if not ref:
        out["note"] = "oracle/_ref (the reference C) is not present here: the oracle's restatement was timed"

    # -- 2048^2 x 1500, in full
    img = O.phantom_fromdesc(random_ellipses(32, 2), n).astype(np.float32).astype(np.float64)
    radon(img[:64, :64].copy(), 64, 8)  # load, first-touch
    t0 = time.perf_counter()
    sino = radon(img, n, angles)
    t1 = time.perf_counter()
    bp(sino, n, angles)
    t2 = time.perf_counter()
    upd = float(n) * n * angles
    tr, tb = t1 - t0, t2 - t1
    out.update({"value": round(2 * upd / (tr + tb) / 1e9, 4),
                "sample": "%dx%d image, all %d angles, %d bins: one Radon and one backprojection, FP64, "
                          "%d threads" % (n, n, angles, n, nthreads),
                "radon_s": round(tr, 3), "backprojection_s": round(tb, 3),
                "radon_gups": round(upd / tr / 1e9, 4), "backprojection_gups": round(upd / tb / 1e9, 4)})

    # -- the reference's 256^2 Shepp-Logan case
    n1, R1, A1 = 256, 256, 180
    from pyraft_amd.phantom import _shepp_logan_desc
    f1 = O.phantom_fromdesc(_shepp_logan_desc(), n1).astype(np.float32).astype(np.float64)
    s1 = radon(f1, R1, A1)
    times = {"radon": [], "backprojection": [], "fbp": []}
    for _ in range(3):
        t0 = time.perf_counter()
        s1 = radon(f1, R1, A1)
        t1 = time.perf_counter()
        bp(s1, n1, A1)
        t2 = time.perf_counter()
        bp(O.lowpassino_fbp(s1), n1, A1)
        t3 = time.perf_counter()
        times["radon"].append(t1 - t0)
        times["backprojection"].append(t2 - t1)
        times["fbp"].append(t3 - t2)
    out["reference_case"] = {"workload": "256x256 Shepp-Logan, 180 angles, 256 bins; median of 3",
                             **{k + "_us": round(statistics.median(v) * 1e6, 1) for k, v in times.items()}}

    # -- one OSEM iteration, 1024^2, 720 angles, 16 subsets
    ne, Ae, S = 1024, 720, 16
    fe = O.phantom_fromdesc(random_ellipses(32, 3), ne)
    se = radon(fe, ne, Ae)
    se = np.random.default_rng(30).poisson(se * (1e3 / se.mean())).astype(np.float64)
    blocks = [((b * Ae) // S, ((b + 1) * Ae) // S) for b in range(S)]
    weights = [bp(np.ones((ne, j1 - j0)), ne, Ae, j0) for j0, j1 in blocks]
    x = np.ones((ne, ne))
    t0 = time.perf_counter()
    for (j0, j1), w in zip(blocks, weights):
        p = radon(x, ne, Ae, j0, j1)
        rho = np.where(p > 0, se[:, j0:j1] / np.where(p > 0, p, 1.0), 0.0)
        x = np.where(w > 0, x * bp(rho, ne, Ae, j0) / np.where(w > 0, w, 1.0), x)
    t_it = time.perf_counter() - t0
    out["osem"] = {"workload": "OSEM 1024x1024, 720 angles, 16 subsets: one iteration (weights precomputed)",
                   "s_per_iteration": round(t_it, 3), "iters_per_s": round(1.0 / t_it, 4),
                   "gups": round(2.0 * ne * ne * Ae / t_it / 1e9, 4)}

    # -- the mesh Radon (non-uniform t and theta tables), on 24 of the mesh leg's angles
    t_m, th_m = _mesh_tables(n, angles)
    th_s = th_m[::max(1, angles // 24)][:24].copy()
    if ref:  # the reference's withmesh computes every ray when nthreads divides R x A
        mesh_radon = lambda: O.ref_radon_withmesh(img, t_m, th_s, nthreads=nthreads)  # noqa: E731
    else:
        mesh_radon = lambda: O.radon_withmesh(img, t_m, th_s, nthreads=nthreads)  # noqa: E731
    t0 = time.perf_counter()
    mesh_radon()
    t_mesh = time.perf_counter() - t0
    out["mesh"] = {"workload": "%dx%d image, %d detector positions x 24 of the mesh leg's angles" % (n, n, n),
                   "radon_s": round(t_mesh, 4), "radon_gups": round(float(n) * n * 24 / t_mesh / 1e9, 4)}

    # -- the line-traversal pair: the reference's own iterator (raft/raft_ddaiterator.c,
    # driven by oracle/ref_dda_harness.c), 24 angles over [0, pi]
    if ref:
        dda_r = lambda: O.ref_radon_bresenham(img, n, 24)  # noqa: E731
        dda_t = lambda s_: O.ref_radon_transpose_bresenham(s_, (n, n))  # noqa: E731
    else:
        dda_r = lambda: O.radon_bresenham(img, n, 24, nthreads=nthreads)  # noqa: E731
        dda_t = lambda s_: O.radon_transpose_bresenham(s_, (n, n), nthreads=nthreads)  # noqa: E731
    t0 = time.perf_counter()
    s24 = dda_r()
    t1 = time.perf_counter()
    dda_t(s24)
    t2 = time.perf_counter()
    upd24 = float(n) * n * 24
    out["bresenham"] = {"workload": "%dx%d image, %d rays x 24 angles over [0, pi]" % (n, n, n),
                        "radon_s": round(t1 - t0, 4), "transpose_s": round(t2 - t1, 4),
                        "radon_gups": round(upd24 / (t1 - t0) / 1e9, 4),
                        "transpose_gups": round(upd24 / (t2 - t1) / 1e9, 4)}

    # -- binary image files: the reference's raft_image_write / raft_image_read
    if ref:
        import tempfile
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "image.raft")
            O.ref_image_write(path, img, (-1.0, 1.0), (1.0, -1.0))  # file created, pages cached
            t0 = time.perf_counter()
            O.ref_image_write(path, img, (-1.0, 1.0), (1.0, -1.0))
            t1 = time.perf_counter()
            O.ref_image_read(path, cap=img.size)
            t2 = time.perf_counter()
        mb = img.size * 8 / 1e6
        out["binary_io"] = {"workload": "%dx%d FP64 image to a file and back" % (n, n),
                            "write_mb_s": round(mb / (t1 - t0), 1), "read_mb_s": round(mb / (t2 - t1), 1)}

    # -- one transmission-EM iteration, 1024^2, 720 angles (raft/raft_transmission_em.cu:41-59
    # composed from the C operators: W = BP(raw) beforehand)
    nt, At = 1024, 720
    xt = O.phantom_fromdesc(random_ellipses(32, 6), nt)
    pt = radon(xt, nt, At)
    xt = xt * (2.0 / pt.max())
    flat_t = np.full((nt, At), 1e4)
    raw_t = flat_t * np.exp(-radon(xt, nt, At))
    w_t = bp(raw_t, nt, At)
    x = np.ones((nt, nt))
    t0 = time.perf_counter()
    r_t = np.where(w_t != 0, bp(flat_t * np.exp(-radon(x, nt, At)), nt, At) / np.where(w_t != 0, w_t, 1.0), 1.0)
    x = (x * r_t) * r_t
    t_tr = time.perf_counter() - t0
    out["transmission_em"] = {"workload": "transmission EM 1024x1024, 720 angles: one iteration (W precomputed)",
                              "s_per_iteration": round(t_tr, 3), "iters_per_s": round(1.0 / t_tr, 4),
                              "gups": round(2.0 * nt * nt * At / t_tr / 1e9, 4)}

    # -- the 256-slice uneven batch, stratified sample, extrapolated
    rng = np.random.default_rng(4)
    sizes = [int(v) for v in rng.choice([256, 512, 768, 1024, 1536, 2048], 256)]
    angs = [int(v) for v in rng.integers(180, 1801, 256)]
    sub = 24
    rate = {}
    picked = []
    for size in sorted(set(sizes)):
        idx = [i for i, v in enumerate(sizes) if v == size]
        for i in sorted({min(idx, key=lambda k: angs[k]), max(idx, key=lambda k: angs[k])}):
            picked.append(i)
            fi = O.phantom_fromdesc(random_ellipses(32, 1000 + i), size)
            t0 = time.perf_counter()
            si = radon(fi, size, sub)
            bp(si, size, sub)
            dt = time.perf_counter() - t0
            rate.setdefault(size, []).append(2.0 * size * size * sub / dt)
    rate = {k: statistics.mean(v) for k, v in rate.items()}
    t_pass = sum(2.0 * s * s * a / rate[s] for s, a in zip(sizes, angs))
    out["batched"] = {"workload": "256 uneven slices (seed 4): Radon + backprojection pass and OSEM 8 subsets x 20 "
                                  "iterations, extrapolated",
                      "sample": "%d slices (per image side the fewest and the most angles), %d angles each; "
                                "per-update rate per image side" % (len(picked), sub),
                      "radon_bp_s_extrapolated": round(t_pass, 2),
                      "osem_iters_per_s_extrapolated": round(1.0 / t_pass, 5),
                      "gups_by_n": {str(k): round(v / 1e9, 4) for k, v in sorted(rate.items())}}
    return out


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
    # warm every plan (its subsets' launch choices are tuned on first use), so the
    # timed loop runs what a user's second call would
    warmed = set()
    for i, (n, a) in enumerate(zip(sizes, angles)):
        if (n, a) not in warmed:
            plans[(n, a)].em(plain_sinos[i], 1, nsubsets)
            warmed.add((n, a))
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


def _mesh_tables(R, A, seed=5):
    """Detector positions and angles of the mesh leg: each sorted uniform random, t in
    [-1, 1] and theta in [0, pi)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    return np.sort(rng.uniform(-1.0, 1.0, R)), np.sort(rng.uniform(0.0, math.pi, A))


def mesh_leg(dev, img, n, R, A, reps=5):
    """Slant-stack Radon on non-uniform (t, theta) tables (raft_radon_slantstack_withmesh,
    pyraft's radon_withmesh) on the headline image: event-timed, plan built beforehand."""
    print("bench: mesh Radon", file=sys.stderr, flush=True)
    import torch
    from pyraft_amd.device import Plan
    t, theta = _mesh_tables(R, A)
    plan = Plan.mesh(n, t, theta)
    plan.radon(img)
    torch.cuda.synchronize()
    ms = _event_ms(lambda: plan.radon(img), reps, torch.cuda.current_stream())
    upd = float(n) * R * A
    return {"workload": "%d^2 image, %d detector positions x %d angles, each sorted uniform random "
                        "(t in [-1, 1], theta in [0, pi))" % (n, R, A),
            "radon_ms": round(ms, 4), "radon_gups": round(upd / ms / 1e6, 1)}


def _tr_em_case(plan, n, dev):
    """Transmission data for the transmission-EM leg: raw = flat exp(-R x*) for a
    32-ellipse attenuation phantom x* scaled so that its longest line integral is 2, flat
    field 1e4."""
    import torch
    from pyraft_amd.device import phantom
    from pyraft_amd.phantom import random_ellipses
    xs = phantom(random_ellipses(32, 6), n)
    xs = xs * (2.0 / float(plan.radon(xs).max()))
    flat = torch.full((plan.nrays, plan.nangles), 1e4, device=dev)
    return flat * torch.exp(-plan.radon(xs)), flat


def tr_em_leg(dev, n=1024, A=720, niter=20):
    """Transmission EM (raft_tr_em_gpu's recurrence, raft/raft_transmission_em.cu:16-66,
    through raft_ss_tr_em on device buffers): iterations per second from x = 1, the fixed
    part of a call (W = BP(raw)) subtracted as in the OSEM leg."""
    print("bench: transmission EM %d^2, %d angles" % (n, A), file=sys.stderr, flush=True)
    import ctypes
    import torch
    from pyraft_amd.device import Plan
    from pyraft_amd.raftypes import check, libraft
    plan = Plan(n, n, A)
    raw, flat = _tr_em_case(plan, n, dev)
    x = torch.ones((n, n), device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def run(k):
        x.fill_(1.0)
        check(libraft.raft_ss_tr_em(plan.handle, raw.data_ptr(), flat.data_ptr(), x.data_ptr(), k, 1, stream),
              "raft_ss_tr_em")

    run(1)  # workspace, kernels
    torch.cuda.synchronize()
    ta = time.perf_counter()
    run(0)
    torch.cuda.synchronize()
    tb = time.perf_counter()
    run(niter)
    torch.cuda.synchronize()
    tc = time.perf_counter()
    it_s = niter / max(1e-9, (tc - tb) - (tb - ta))
    return {"workload": "transmission EM %dx%d, %d angles, %d bins, %d iterations from x = 1; raw = 1e4 "
                        "exp(-R x*), line integrals up to 2" % (n, n, A, n, niter),
            "iters_per_s": round(it_s, 2), "gups": round(2.0 * n * n * A * it_s / 1e9, 2)}


def io_leg(n=2048):
    """Binary raft image files (raft_image_write / raft_image_read, raft/raft_image.c:5-38,
    135-151): an n x n FP64 image written to a temporary file and read back (host only)."""
    print("bench: binary image files", file=sys.stderr, flush=True)
    import tempfile
    import numpy as np
    import pyraft_amd as pr
    img = pr.image(np.random.default_rng(7).random((n, n)))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "image.raft")
        pr.image_write(img, path)  # file created, pages cached
        t0 = time.perf_counter()
        pr.image_write(img, path)
        t1 = time.perf_counter()
        back = pr.image_read(path)
        t2 = time.perf_counter()
    mb = n * n * 8 / 1e6
    return {"workload": "%dx%d FP64 image to a file and back" % (n, n),
            "write_mb_s": round(mb / (t1 - t0), 1), "read_mb_s": round(mb / (t2 - t1), 1),
            "round_trip_exact": bool(np.array_equal(np.asarray(back), np.asarray(img)))}


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: measures one GPU only (got --gpus %d, WORLD_SIZE %d); "
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
    rt_variant, rt_variant_us = plan.tune_transpose(nslices=1, reps=3)  # and the exact transpose's
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
    traffic, traffic_src = measured_traffic(dominant)
    roofline = roofline_of(dominant, upd, t_dom, traffic, traffic_src, measured_counters(dominant))
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
        "config": {"workload": "slant-stack Radon + backprojection, %dx%d image, %d angles in [0,pi), %d bins; "
                               "device-resident, angle-major sinogram between the two (the (rays, angles) entry "
                               "adds one %d x %d transpose)" % (n, n, A, R, R, A), "n": n, "angles": A, "rays": R},
        "radon_gups": round(upd / t_radon / 1e9, 2),
        "backprojection_gups": round(upd / t_bp / 1e9, 2),
        "radon_ms": round(t_radon * 1e3, 4),
        "backprojection_ms": round(t_bp * 1e3, 4),
        "fbp_ms": round(fbp_ms, 4),
        "exact_transpose_ms": round(transpose_ms, 4),
        "bp_variant": {"chosen": bp_variant, "rows_per_thread": bp_rows, "stage_pairs": bp_stage,
                       "tuned_us": {str(k): round(v, 1) for k, v in bp_variant_us.items()}},
        "transpose_variant": {"chosen": rt_variant, "tuned_us": {str(k): round(v, 1) for k, v in rt_variant_us.items()}},
        "roofline": roofline,
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

    if not args.no_rows:
        result["mesh"] = mesh_leg(dev, img, n, R, A)
        result["transmission_em"] = tr_em_leg(dev)
        result["binary_io"] = io_leg()

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
    aplan.tune_transpose(nslices=64, reps=2)   # the backward's tile shape for 64-slice launches
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
