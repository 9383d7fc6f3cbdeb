"""GPU parity of the HIP slant-stack path against the FP64 oracle.

Inputs are FP32 arrays; the oracle receives the same values cast to FP64.  Bars:
relative L2 <= 1e-5 for Radon, backprojection, FBP and EM,
adjoint mismatch <= 1e-6 for the exact transpose, bit-exact plan ranges.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5
ADJOINT = 1e-6


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


def _plan(n, R, A, ic=(-1.0, 1.0, 1.0, -1.0), sc=(0.0, 1.0, math.pi, -1.0)):
    from pyraft_amd.device import Plan
    return Plan(n, R, A, ic, sc)


def _geo(ic, sc):
    return dict(img_tl=(ic[0], ic[1]), img_br=(ic[2], ic[3]), sino_tl=(sc[0], sc[1]), sino_br=(sc[2], sc[3]))


def _ellipse_image(n, seed, count=32):
    from pyraft_amd.phantom import random_ellipses
    desc = random_ellipses(count, seed)
    return O.phantom_fromdesc(desc, n).astype(np.float32)


GEOMS = [
    # n, R, A, img corners, sino corners
    (64, 64, 45, (-1.0, 1.0, 1.0, -1.0), (0.0, 1.0, math.pi, -1.0)),
    (65, 70, 37, (-1.0, 1.0, 1.0, -1.0), (0.0, 1.0, math.pi, -1.0)),
    (48, 40, 91, (-0.7, 1.3, 1.2, -0.9), (-math.pi / 2, 1.1, math.pi / 2, -0.8)),
    (256, 256, 180, (-1.0, 1.0, 1.0, -1.0), (0.0, 1.0, math.pi, -1.0)),
]


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: "n%d_R%d_A%d" % g[:3])
def test_plan_bounds_match_host(torch_mod, geom):
    import ctypes
    from pyraft_amd.raftypes import libraft
    n, R, A, ic, sc = geom
    print("plan", geom, flush=True)
    p = _plan(n, R, A, ic, sc)
    dev = p.bounds()
    host = np.zeros((A, R, 2), np.int32)
    rc = libraft.raft_ss_plan_host_bounds(n, R, A, (ctypes.c_double * 4)(*ic), (ctypes.c_double * 4)(*sc),
                                          host.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None, None)
    assert rc == 0
    assert np.array_equal(dev, host)
    # per 64-ray group: union / intersection over the non-empty rays
    wb = p.wave_bounds()
    for j in range(A):
        for g in range((R + 63) // 64):
            rows = host[j, g * 64:(g + 1) * 64]
            ne = rows[:, 1] >= rows[:, 0]
            if not ne.any():
                expect = (0, -1, 0, -1)
            else:
                expect = (rows[ne, 0].min(), rows[ne, 1].max(), rows[ne, 0].max(), rows[ne, 1].min())
            assert tuple(wb[j, g]) == tuple(expect), (j, g)


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: "n%d_R%d_A%d" % g[:3])
def test_radon_parity(torch_mod, geom):
    torch = torch_mod
    n, R, A, ic, sc = geom
    print("radon", geom, flush=True)
    rng = np.random.default_rng(11)
    x = rng.random((n, n)).astype(np.float32)
    p = _plan(n, R, A, ic, sc)
    got = p.radon(torch.from_numpy(x).cuda()).cpu().numpy()
    ref = O.radon(x.astype(np.float64), R, A, **_geo(ic, sc))
    err = rel_l2(got, ref)
    print("rel-L2", err)
    assert err <= REL_L2
    # angle-major output is the transpose
    gotT = p.radon(torch.from_numpy(x).cuda(), angle_major=True).cpu().numpy()
    assert np.array_equal(gotT, got.T)


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: "n%d_R%d_A%d" % g[:3])
def test_backprojection_parity(torch_mod, geom):
    torch = torch_mod
    n, R, A, ic, sc = geom
    print("bp", geom, flush=True)
    rng = np.random.default_rng(12)
    s = rng.random((R, A)).astype(np.float32)
    p = _plan(n, R, A, ic, sc)
    got = p.backprojection(torch.from_numpy(s).cuda()).cpu().numpy()
    ref = O.backprojection(s.astype(np.float64), n, **_geo(ic, sc))
    err = rel_l2(got, ref)
    print("rel-L2", err)
    assert err <= REL_L2


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: "n%d_R%d_A%d" % g[:3])
def test_transpose_adjoint(torch_mod, geom):
    torch = torch_mod
    n, R, A, ic, sc = geom
    print("transpose", geom, flush=True)
    rng = np.random.default_rng(13)
    x = rng.random((n, n)).astype(np.float32)
    y = rng.random((R, A)).astype(np.float32)
    p = _plan(n, R, A, ic, sc)
    Rx = p.radon(torch.from_numpy(x).cuda()).cpu().numpy().astype(np.float64)
    Rty = p.radon_transpose(torch.from_numpy(y).cuda()).cpu().numpy().astype(np.float64)
    lhs = float(np.sum(Rx * y.astype(np.float64)))
    rhs = float(np.sum(x.astype(np.float64) * Rty))
    mism = abs(lhs - rhs) / abs(lhs)
    print("adjoint mismatch", mism)
    assert mism <= ADJOINT
    ref = O.radon_transpose(y.astype(np.float64), n, **_geo(ic, sc))
    assert rel_l2(Rty, ref) <= REL_L2


DENSE_GEOMS = [
    # rays 0.32-0.45 cells apart: lanes 3-4 apart update the line in turn
    (48, 150, 60, (-1.0, 1.0, 1.0, -1.0), (0.0, 1.0, math.pi, -1.0)),
    # rays 0.08-0.11 cells apart: below 1/8 cell the LDS atomic form
    (32, 400, 40, (-1.0, 1.0, 1.0, -1.0), (0.0, 1.0, math.pi, -1.0)),
    # rays 0.9-1.3 cells apart over 2 pi, long lines: interior passes without range reads
    (300, 200, 90, (-1.0, 1.0, 1.0, -1.0), (0.0, 0.6, 2 * math.pi, -0.6)),
]


@pytest.mark.parametrize("geom", DENSE_GEOMS, ids=lambda g: "n%d_R%d_A%d" % g[:3])
def test_transpose_dense_rays_and_interior_passes(torch_mod, geom):
    torch = torch_mod
    n, R, A, ic, sc = geom
    print("transpose", geom, flush=True)
    rng = np.random.default_rng(29)
    x = rng.random((n, n)).astype(np.float32)
    y = rng.random((R, A)).astype(np.float32)
    p = _plan(n, R, A, ic, sc)
    Rx = p.radon(torch.from_numpy(x).cuda()).cpu().numpy().astype(np.float64)
    Rty = p.radon_transpose(torch.from_numpy(y).cuda()).cpu().numpy().astype(np.float64)
    lhs = float(np.sum(Rx * y.astype(np.float64)))
    rhs = float(np.sum(x.astype(np.float64) * Rty))
    mism = abs(lhs - rhs) / abs(lhs)
    print("adjoint mismatch", mism)
    assert mism <= ADJOINT
    ref = O.radon_transpose(y.astype(np.float64), n, **_geo(ic, sc))
    assert rel_l2(Rty, ref) <= REL_L2


def test_radon_random_ellipses_512(torch_mod):
    torch = torch_mod
    n = R = 512
    A = 360
    print("ellipses 512", flush=True)
    x = _ellipse_image(n, 2)
    p = _plan(n, R, A)
    got = p.radon(torch.from_numpy(x).cuda()).cpu().numpy()
    ref = O.radon(x.astype(np.float64), R, A)
    err = rel_l2(got, ref)
    print("radon rel-L2", err)
    assert err <= REL_L2
    bp = p.backprojection(torch.from_numpy(ref.astype(np.float32)).cuda()).cpu().numpy()
    bref = O.backprojection(ref.astype(np.float32).astype(np.float64), n)
    err = rel_l2(bp, bref)
    print("bp rel-L2", err)
    assert err <= REL_L2


def test_fbp_filter_strided_layouts(torch_mod):
    """Small sinograms are filtered straight from the (R, A) layout (strided rows, two
    rows per block, which cross from one slice to the next when A is odd) and, for the
    filter alone, written back to it; larger ones go through angle-major copies.  Both
    against the oracle per slice, with several slices per launch."""
    torch = torch_mod
    n, R, A, ns = 96, 101, 37, 3  # odd A: the last row pair of a slice spans two slices
    p = _plan(n, R, A)
    rng = np.random.default_rng(5)
    s = rng.random((ns, R, A)).astype(np.float32)
    st = torch.from_numpy(s).cuda()
    filt = p.fbp_filter(st).cpu().numpy()
    filt_am = p.fbp_filter(st.transpose(-1, -2).contiguous(), angle_major=True).transpose(-1, -2).cpu().numpy()
    rec = p.fbp(st).cpu().numpy()
    for i in range(ns):
        s64 = s[i].astype(np.float64)
        assert rel_l2(filt[i], O.lowpassino_fbp(s64)) <= REL_L2, i
        assert rel_l2(filt_am[i], O.lowpassino_fbp(s64)) <= REL_L2, i
        assert rel_l2(rec[i], O.fbp(s64, n)) <= REL_L2, i


def test_fbp_parity(torch_mod):
    torch = torch_mod
    n = R = 256
    A = 180
    print("fbp", flush=True)
    x = _ellipse_image(n, 7)
    s = O.radon(x.astype(np.float64), R, A).astype(np.float32)
    p = _plan(n, R, A)
    filt = p.fbp_filter(torch.from_numpy(s).cuda()).cpu().numpy()
    fref = O.lowpassino_fbp(s.astype(np.float64))
    assert rel_l2(filt, fref) <= REL_L2
    got = p.fbp(torch.from_numpy(s).cuda()).cpu().numpy()
    ref = O.fbp(s.astype(np.float64), n)
    err = rel_l2(got, ref)
    print("fbp rel-L2", err)
    assert err <= REL_L2


@pytest.mark.parametrize("nsubsets", [1, 3])
def test_em_parity(torch_mod, nsubsets):
    torch = torch_mod
    n = R = 96
    A = 60
    niter = 4
    print("em", nsubsets, flush=True)
    x = _ellipse_image(n, 3, count=12)
    p_true = O.radon(x.astype(np.float64), R, A)
    s = (p_true * 50.0).astype(np.float32)
    p = _plan(n, R, A)
    got, ll = p.em(torch.from_numpy(s).cuda(), niter, nsubsets, return_ll=True)
    got = got.cpu().numpy()
    ref, llref = O.osem(s.astype(np.float64), n, niter, nsubsets)
    err = rel_l2(got, ref)
    print("em rel-L2", err, "ll", ll, llref)
    assert err <= REL_L2
    for a, b in zip(ll, llref):
        assert abs(a - b) <= 1e-5 * abs(b)


def test_phantom_device(torch_mod):
    from pyraft_amd.device import phantom
    from pyraft_amd.phantom import random_ellipses
    desc = random_ellipses(32, 5)
    got = phantom(desc, 200).cpu().numpy()
    ref = O.phantom_fromdesc(desc, 200)
    assert np.array_equal(got, ref.astype(np.float32))


def test_transpose_wide_image(torch_mod):
    """A wide image (3000-pixel lines in LDS, the ray ranges and sinogram values read
    from global memory): same exactness bars (adjoint identity, oracle)."""
    torch = torch_mod
    n = R = 3000
    A = 24
    print("transpose wide image", n, R, A, flush=True)
    rng = np.random.default_rng(17)
    x = rng.random((n, n)).astype(np.float32)
    y = rng.random((R, A)).astype(np.float32)
    p = _plan(n, R, A)
    Rx = p.radon(torch.from_numpy(x).cuda()).cpu().numpy().astype(np.float64)
    Rty = p.radon_transpose(torch.from_numpy(y).cuda()).cpu().numpy().astype(np.float64)
    lhs = float(np.sum(Rx * y.astype(np.float64)))
    rhs = float(np.sum(x.astype(np.float64) * Rty))
    mism = abs(lhs - rhs) / abs(lhs)
    print("adjoint mismatch", mism)
    assert mism <= ADJOINT
    ref = O.radon_transpose(y.astype(np.float64), n)
    assert rel_l2(Rty, ref) <= REL_L2


@pytest.mark.parametrize("nslices", [2, 3, 5])
@pytest.mark.parametrize("R", [100, 90])
def test_transpose_slice_batches(torch_mod, nslices, R):
    """Slices of a batch go two at a time (an odd last one alone): every slice equals the
    single-slice transpose bit for bit (the same operations in the same order).  R = 100 >
    n puts rays closer than a pixel (the line form, two slices per wave); R = 90 keeps
    them at least a pixel apart (the pixel form, two slices per block)."""
    torch = torch_mod
    n, A = 96, 37
    print("transpose slices", nslices, "R", R, flush=True)
    p = _plan(n, R, A)
    y = torch.from_numpy(np.random.default_rng(nslices).random((nslices, R, A)).astype(np.float32)).cuda()
    got = p.radon_transpose(y).cpu().numpy()
    for i in range(nslices):
        one = p.radon_transpose(y[i]).cpu().numpy()
        assert np.array_equal(got[i], one), i
    assert rel_l2(got[-1], O.radon_transpose(y[-1].cpu().numpy().astype(np.float64), n)) <= REL_L2


def test_transpose_angle_ranges(torch_mod):
    """The transpose of an angle range [j0, j1) (the autograd backward of a range
    forward) from both layouts and with slice pairs plus a single slice: the oracle's
    transpose of the sinogram with the other angles zeroed.  The launch pre-scales only
    the range's rows and interleaves the pair, so a wrong row or slice offset fails."""
    torch = torch_mod
    n, R, A, ns = 128, 120, 90, 3
    p = _plan(n, R, A)
    s = np.random.default_rng(31).random((ns, R, A)).astype(np.float32)
    st = torch.from_numpy(s).cuda()
    for j0, j1 in ((0, A), (17, 61), (89, 90)):
        print("transpose range", j0, j1, flush=True)
        got = p.radon_transpose(st, angles=(j0, j1)).cpu().numpy()
        got_am = p.radon_transpose(st.transpose(-1, -2).contiguous(), angle_major=True, angles=(j0, j1)).cpu().numpy()
        for i in range(ns):
            y = np.zeros((R, A))
            y[:, j0:j1] = s[i, :, j0:j1]
            ref = O.radon_transpose(y, n)
            assert rel_l2(got[i], ref) <= REL_L2, (j0, j1, i)
            assert np.array_equal(got_am[i], got[i]), (j0, j1, i)


@pytest.mark.parametrize("n,R,A,ns", [(256, 256, 180, 1), (100, 103, 61, 3), (64, 64, 200, 2)])
def test_small_image_angle_groups(torch_mod, n, R, A, ns):
    """Small images divide a backprojection's angles into groups (partial images added
    in group order by a finishing pass) so that the launch fills the GPU: plain
    backprojection, FBP and EM (weights and multiplicative update) against the oracle,
    several slices per launch, every slice as its own one-slice call."""
    torch = torch_mod
    p = _plan(n, R, A)
    print("angle groups n=%d R=%d A=%d slices %d: %d groups" % (n, R, A, ns, p.info()["bp_groups"]), flush=True)
    assert p.info()["bp_groups"] > 1
    rng = np.random.default_rng(n + A)
    s = rng.random((ns, R, A)).astype(np.float32)
    st = torch.from_numpy(s).cuda()
    got = p.backprojection(st).cpu().numpy()
    fbp = p.fbp(st).cpu().numpy()
    for i in range(ns):
        s64 = s[i].astype(np.float64)
        assert rel_l2(got[i], O.backprojection(s64, n)) <= REL_L2, i
        assert rel_l2(fbp[i], O.fbp(s64, n)) <= REL_L2, i
        one = p.backprojection(st[i]).cpu().numpy()
        assert rel_l2(one, got[i]) <= 1e-6, i
    x = (0.5 + rng.random((n, n))).astype(np.float32)
    meas = p.radon(torch.from_numpy(x).cuda()) * 20.0
    em, ll = p.em(meas, 2, 4, return_ll=True)
    ref, llref = O.osem(meas.cpu().numpy().astype(np.float64), n, 2, 4)
    e = rel_l2(em.cpu().numpy(), ref)
    print("  em rel-L2 %.2e" % e, flush=True)
    assert e <= REL_L2
    for a, b in zip(ll, llref):
        assert abs(a - b) <= 1e-6 * abs(b)
