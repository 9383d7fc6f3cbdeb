"""The pyraft-compatible entry points on the GPU against the FP64 oracle.

Covers the reference's CPU-runnable workload (256 x 256 Shepp-Logan, 180 angles,
256 bins: Radon, backprojection, FBP) and every drop-in symbol: radon/backprojection
through raft_image structs (including strided views), radon_gpu, backprojection_gpu,
fbp_gpu, em_gpu, lowpassino_fbp, radon_transpose, make_em (slant-stack) / osem.
Tolerance: relative L2 <= 1e-5 against the oracle on the same (FP32-rounded) inputs.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
REL_L2 = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def pr():
    import torch
    assert torch.cuda.is_available()
    import pyraft_amd
    return pyraft_amd


def _shepp_logan_table():
    import importlib
    return importlib.import_module("pyraft_amd.phantom")._shepp_logan_desc()


def _f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def test_shepp_logan_256_radon_bp_fbp(pr):
    print("C1 shepp-logan", flush=True)
    f = pr.shepplogan((256, 256))
    assert np.array_equal(f, O.phantom_fromdesc(_shepp_logan_table(), 256))
    s = pr.radon(f, (256, 180), method=pr.radon_method.SLANTSTACK)
    assert s.top_left == (0.0, 1.0) and s.bottom_right == (math.pi, -1.0)
    s_ref = O.radon(_f32(f), 256, 180)
    e = rel_l2(s, s_ref)
    print("radon", e)
    assert e <= REL_L2
    plan = pr.backprojection_plan(s.shape, (256, 256), method=pr.radon_method.SLANTSTACK)
    b = pr.backprojection(s, plan)
    e = rel_l2(b, O.backprojection(_f32(s), 256))
    print("backprojection", e)
    assert e <= REL_L2
    rec = pr.fbp(s)
    e = rel_l2(rec, O.fbp(_f32(s), 256))
    print("fbp", e)
    assert e <= REL_L2
    filt = pr.lowpassino_fbp(s)
    assert filt.top_left == s.top_left
    assert rel_l2(filt, O.lowpassino_fbp(_f32(s))) <= REL_L2


def test_strided_and_flipped_views(pr):
    rng = np.random.default_rng(3)
    base = rng.random((96, 96))
    view = pr.image(base)[::-1, :]  # negative column stride through raft_matrix
    view.top_left, view.bottom_right = (-1.0, 1.0), (1.0, -1.0)
    s = pr.radon(view, (96, 40), method=pr.radon_method.SLANTSTACK)
    assert rel_l2(s, O.radon(_f32(np.ascontiguousarray(view)), 96, 40)) <= REL_L2
    sF = np.asfortranarray(np.asarray(s))  # column-major sinogram (line_stride = R)
    assert not sF.flags.c_contiguous
    b = pr.backprojection(sF, (pr.radon_method.SLANTSTACK, (96, 96), 2))
    assert rel_l2(b, O.backprojection(_f32(s), 96)) <= REL_L2


def test_legacy_float_entry_points(pr):
    n, R, A = 128, 128, 90
    f = pr.ellipses_phantom(pr.random_ellipses(20, 4), (n, n))
    s = pr.radon_gpu(np.asarray(f), R, A)
    assert s.top_left == (0, 1.0) and s.bottom_right == (math.pi, -1.0)
    assert rel_l2(s, O.radon(_f32(f), R, A)) <= REL_L2
    b = pr.backprojection_gpu(np.asarray(s), (n, n))
    assert b.dtype == np.float64
    assert rel_l2(b, O.backprojection(_f32(s), n)) <= REL_L2
    fb = pr.fbp_gpu(s)
    assert fb.dtype == np.float32 and fb.shape == (n, n)
    assert rel_l2(fb, O.fbp(_f32(s), n)) <= REL_L2
    em = pr.em_gpu(np.asarray(s) * 100.0, n, 4)
    assert em.dtype == np.float32
    ref, _ = O.em(_f32(np.asarray(s) * 100.0), n, 4)
    assert rel_l2(em, ref) <= REL_L2


def test_radon_transpose_and_fromdesc(pr):
    n, R, A = 80, 90, 50
    rng = np.random.default_rng(8)
    y = pr.image(rng.random((R, A)), x_extent=(0.0, math.pi))
    g = pr.radon_transpose(y, (n, n), method=pr.radon_method.SLANTSTACK)
    assert rel_l2(g, O.radon_transpose(_f32(y), n)) <= REL_L2
    d = pr.random_ellipses(8, 6)
    exact = pr.radon(d, (R, A))
    assert rel_l2(exact, O.radon_fromdesc(d, R, A)) <= 1e-12


def test_em_osem_make_em(pr):
    n = R = 64
    A = 48
    f = pr.ellipses_phantom(pr.random_ellipses(10, 9), (n, n))
    s = pr.radon(f, (R, A), method=pr.radon_method.SLANTSTACK) * 40.0
    # the fused slant-stack iteration (pyraft's em() defaults to the Bresenham pair,
    # covered in test_bresenham.py)
    ss = pr.radon_method.SLANTSTACK
    step = pr.make_em(s, (n, n), projection_method=ss, backprojection_method=ss)
    x1 = step(pr.image(np.ones((n, n))))
    ref1, _ = O.em(_f32(s), n, 1)
    assert rel_l2(x1, ref1) <= REL_L2
    x = x1
    for _ in range(2):
        x = step(x)
    ref, _ = O.em(_f32(s), n, 3)
    assert rel_l2(x, ref) <= REL_L2
    xo = pr.osem(s, niter=2, nblocks=4)
    refo, _ = O.osem(_f32(s), n, 2, nsubsets=4)
    assert rel_l2(xo, refo) <= REL_L2


def test_large_host_arrays_match_device_path(pr):
    """Images of 2^20 elements and more convert FP64 <-> FP32 on several host threads and
    stage through pooled pinned buffers: the host-array drop-ins must give exactly what the
    device-resident calls give on the same FP32-rounded data (a strided view included, and
    two calls in a row reusing the pooled buffers)."""
    import torch
    from pyraft_amd.device import Plan
    n, R, A = 1100, 1100, 64
    print("large host arrays", n, R, A, flush=True)
    rng = np.random.default_rng(1100)
    view = pr.image(rng.random((2 * n, 2 * n)))[::2, ::2]   # strides of 2 rows and 2 columns
    view.top_left, view.bottom_right = (-1.0, 1.0), (1.0, -1.0)
    assert not view.flags.c_contiguous
    p = Plan(n, R, A)
    ref_sino = p.radon(torch.from_numpy(np.asarray(view, dtype=np.float32)).cuda()).double().cpu().numpy()
    for _ in range(2):
        s = pr.radon(view, (R, A), method=pr.radon_method.SLANTSTACK)
        assert np.array_equal(np.asarray(s), ref_sino)
    y = np.ascontiguousarray(ref_sino)
    ref_bp = p.backprojection(torch.from_numpy(y.astype(np.float32)).cuda()).double().cpu().numpy()
    plan = pr.backprojection_plan(y.shape, (n, n), method=pr.radon_method.SLANTSTACK)
    for _ in range(2):
        b = pr.backprojection(y, plan)
        assert np.array_equal(np.asarray(b), ref_bp)


def test_radon_out_is_validated(pr):
    """Plan.radon(out=...) hands out's data pointer to the kernel: a tensor of another
    shape, dtype or layout is refused instead of written past its end."""
    import torch
    from pyraft_amd.device import Plan
    print("radon out validation", flush=True)
    n, R, A = 64, 64, 30
    plan = Plan(n, R, A)
    x = torch.rand((n, n), device="cuda")
    ok = torch.empty((A, R), device="cuda")
    plan.radon(x, angle_major=True, out=ok)
    for bad, exc in ((torch.empty((A - 1, R), device="cuda"), ValueError),
                     (torch.empty((R, A), device="cuda"), ValueError),
                     (torch.empty((A, R), device="cuda", dtype=torch.float64), TypeError),
                     (torch.empty((R, A), device="cuda").t(), ValueError),
                     (torch.empty((A, R)), TypeError)):
        with pytest.raises(exc):
            plan.radon(x, angle_major=True, out=bad)
    ref = plan.radon(x, angle_major=True)
    assert torch.equal(ok, ref)


def test_em_under_stream_capture(pr):
    """raft_ss_em without ll_host enqueues kernels and stream-ordered workspace only, so
    it can be captured into a HIP graph: every replay (with eager work and host waits
    between replays) equals the eager call bit for bit.  With ll_host it is refused
    while the stream is being captured."""
    import torch
    from pyraft_amd.device import Plan
    from pyraft_amd.raftypes import RaftError
    print("em under stream capture", flush=True)
    n, A, S, niter = 96, 60, 4, 3
    plan = Plan(n, n, A)
    f = torch.rand((n, n), device="cuda") + 0.5
    s = plan.radon(f) * 50.0
    eager = plan.em(s, niter, S)          # tunes the subsets' variant (cached for the capture)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        with torch.cuda.graph(g, stream=stream):
            xg = plan.em(s, niter, S)
    torch.cuda.current_stream().wait_stream(stream)
    for rep in range(3):
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(xg, eager), rep
        plan.radon(torch.rand((n, n), device="cuda"))   # other work between replays
        torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        g2 = torch.cuda.CUDAGraph()
        with pytest.raises(RaftError):
            with torch.cuda.graph(g2, stream=stream):
                plan.em(s, niter, S, return_ll=True)
    torch.cuda.synchronize()
