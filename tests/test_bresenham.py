"""Line-traversal (Bresenham / DDA) projector pair: raft_radon_bresenham,
raft_radon_transpose_bresenham, raft_backprojection_bresenham (pyraft's default
``radon`` / ``radon_transpose`` / ``em`` method).

CPU (oracle and host logic):
* golden vectors written by the reference's own iterator (raft/raft_ddaiterator.c,
  tests/golden/make_golden.py) are reproduced bit for bit, and so are fresh random
  geometries where oracle/_ref is built;
* for square pixels every visited pixel gets its exact geometric intersection length
  (line clipped to the pixel square), and the Radon of a constant image is the chord
  through the pixel-edge box: a wrong entry pixel, a swapped axis or a lost slow
  move fails these;
* the transpose is the adjoint of the Radon, the backprojection of a constant
  sinogram is that constant wherever a ray passes;
* the library's host build of the device traversal visits exactly the oracle's
  pixels, in order, with the same lengths (1e-13);
* on the angles the transpose evaluates pixel by pixel, every traced length equals
  the pixel's trapezoid in the ray's distance from the pixel centre (1e-12), the
  closed form that transpose uses.

GPU (device API, drop-in symbols, pyraft wrappers): rel-L2 <= 1e-5 against the FP64
oracle (golden and random geometries, rectangular images with non-square pixels,
several slices per launch), adjoint identity to 1e-6, NaN exactly where the
reference divides 0/0, and the default-method ``em`` against the oracle composition.
"""
import ctypes
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "slantstack_golden.npz")
DDA_CASES = ("dda_24x24_r30_a20", "dda_20x31_r25_a17_rect", "dda_16x16_r21_a9_flip")
DEF_IC = (-1.0, 1.0, 1.0, -1.0)
DEF_SC = (0.0, 1.0, math.pi, -1.0)
REL_L2 = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def _geo(ic, sc):
    return dict(img_tl=tuple(ic[:2]), img_br=tuple(ic[2:]), sino_tl=tuple(sc[:2]), sino_br=tuple(sc[2:]))


def _golden_case(golden, c):
    g = golden[c + "/geom"]
    shape, R, A = (int(g[0]), int(g[1])), int(g[2]), int(g[3])
    return shape, R, A, tuple(g[4:8]), tuple(g[8:12])


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


def clip_len(c, s, t, x0, x1, y0, y1):
    """Length of the line {x c + y s = t} inside [x0, x1] x [y0, y1]."""
    lo, hi = -np.inf, np.inf
    for d, a, b, base in ((-s, x0, x1, t * c), (c, y0, y1, t * s)):
        if d == 0:
            if base < a or base > b:
                return 0.0
            continue
        p1, p2 = sorted(((a - base) / d, (b - base) / d))
        lo, hi = max(lo, p1), min(hi, p2)
    return max(0.0, hi - lo)


# (shape, R, A, image corners, sinogram corners)
GEOMS = [
    ((16, 16), 12, 9, DEF_IC, DEF_SC),
    ((33, 33), 40, 31, DEF_IC, DEF_SC),
    ((20, 31), 25, 17, (-0.7, 1.3, 1.2, -0.9), (-math.pi / 2, 1.1, math.pi / 2, -0.8)),   # non-square pixels
    ((24, 24), 30, 45, (1.0, -1.0, -1.0, 1.0), (math.pi, -1.2, 0.0, 1.2)),                # flipped axes
    ((40, 40), 41, 4, DEF_IC, (0.0, 1.0, math.pi / 2, -1.0)),                            # 0, 30, 60, 90 degrees
    ((17, 29), 19, 13, DEF_IC, (0.0, 1.5, 2 * math.pi, -1.5)),                           # rays off the image
    ((64, 20), 70, 33, (-0.5, 2.0, 0.5, -2.0), (0.1, 2.2, 3.0, -2.2)),                  # tall image
]


# ---------------------------------------------------------------- CPU


def test_golden_vectors(golden):
    for c in DDA_CASES:
        shape, R, A, ic, sc = _golden_case(golden, c)
        geo = _geo(ic, sc)
        img, sino = golden[c + "/img"], golden[c + "/sino"]
        assert np.array_equal(O.radon_bresenham(img, R, A, nthreads=3, **geo), golden[c + "/radon"]), c
        assert np.array_equal(O.radon_transpose_bresenham(sino, shape, nthreads=1, **geo), golden[c + "/transpose"]), c
        assert np.array_equal(O.backprojection_bresenham(sino, shape, nthreads=1, **geo), golden[c + "/backprojection"],
                              equal_nan=True), c
        # the reference's thread partition changes only the FP64 summation order
        t4 = O.radon_transpose_bresenham(sino, shape, nthreads=4, **geo)
        assert np.abs(t4 - golden[c + "/transpose"]).max() <= 1e-13 * np.abs(t4).max(), c


@pytest.mark.skipif(not O.ref_available(), reason="reference checker oracle/_ref not built here")
@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: "%dx%d_R%d_A%d" % (g[0] + g[1:3]))
def test_against_reference_iterator(geom):
    shape, R, A, ic, sc = geom
    geo = _geo(ic, sc)
    rng = np.random.default_rng(R * A)
    img, sino = rng.random(shape), rng.random((R, A))
    assert np.array_equal(O.radon_bresenham(img, R, A, nthreads=2, **geo), O.ref_radon_bresenham(img, R, A, **geo))
    assert np.array_equal(O.radon_transpose_bresenham(sino, shape, nthreads=1, **geo),
                          O.ref_radon_transpose_bresenham(sino, shape, **geo))
    assert np.array_equal(O.backprojection_bresenham(sino, shape, nthreads=1, **geo),
                          O.ref_backprojection_bresenham(sino, shape, **geo), equal_nan=True)


@pytest.mark.parametrize("shape,ic", [((16, 16), DEF_IC), ((10, 10), (1.0, -1.0, -1.0, 1.0)),
                                      ((13, 13), (-0.5, 2.0, 1.9, -0.4))])
def test_square_pixel_lengths_are_exact_geometry(shape, ic):
    rng = np.random.default_rng(shape[0])
    L, C = shape
    hx, hy = (ic[2] - ic[0]) / (C - 1), (ic[3] - ic[1]) / (L - 1)
    xc = ic[0] + np.arange(C) * hx
    yc = ic[1] + np.arange(L) * hy
    for trial in range(40):
        th = [0.0, math.pi / 2, math.pi / 4, math.pi, 3 * math.pi / 4][trial] if trial < 5 else rng.uniform(-4, 4)
        t = rng.uniform(-1.3, 1.3) + 0.5 * (ic[0] + ic[2])
        # one ray: a 2 x 2 sinogram whose ray 0 / angle 0 is (theta, t)
        ij, ln = O.dda_ray_path(shape, 2, 2, 0, 0, ic[:2], ic[2:], (th, t), (th + 1.0, t + 1.0))
        geo = np.array([[clip_len(math.cos(th), math.sin(th), t, x - abs(hx) / 2, x + abs(hx) / 2,
                                  y - abs(hy) / 2, y + abs(hy) / 2) for x in xc] for y in yc])
        got = np.zeros(shape)
        np.add.at(got, (ij[:, 0], ij[:, 1]), ln)
        assert np.abs(got - geo).max() <= 1e-12, (th, t)
        assert len({tuple(p) for p in ij}) == len(ij)      # each pixel once


def test_radon_of_constant_is_chord_length():
    n, R, A = 32, 37, 29
    geo = _geo(DEF_IC, DEF_SC)
    s = O.radon_bresenham(np.ones((n, n)), R, A, **geo)
    h = 2.0 / (n - 1)
    box = (-1 - h / 2, 1 + h / 2, -1 - h / 2, 1 + h / 2)
    for a in range(A):
        th = a * (math.pi / (A - 1))
        for r in range(R):
            t = 1.0 + r * (-2.0 / (R - 1))
            assert abs(s[r, a] - clip_len(math.cos(th), math.sin(th), t, *box)) <= 1e-12, (r, a)


def _pixel_driven(c, s, hx, hy):
    """The angles the library's transpose evaluates pixel by pixel: exact geometry
    (slow-axis drift per move within the pixel width) and not within 1e-7 of an axis."""
    fast_i = abs(hx * s) <= abs(hy * c)
    e_step = abs(hy * s / c) if fast_i else abs(hx * c / s)
    e_cap = abs(hx) if fast_i else abs(hy)
    return min(abs(c), abs(s)) >= 1e-7 and e_step <= e_cap


@pytest.mark.parametrize("geom", [GEOMS[0], GEOMS[2], GEOMS[3], ((17, 29), 19, 37, DEF_IC, (0.0, 1.5, 2 * math.pi, -1.5)),
                                  ((30, 17), 33, 60, DEF_IC, DEF_SC)],
                         ids=lambda g: "%dx%d_A%d" % (g[0] + g[2:3]))
def test_trapezoid_lengths_of_pixel_driven_angles(geom):
    """On those angles every traced length is the w x h pixel's trapezoid in the
    distance d = t - (x cos + y sin): Lmax for |d| <= q, linear to 0 at |d| = p."""
    shape, R, A, ic, sc = geom
    L, C = shape
    hx, hy = (ic[2] - ic[0]) / (C - 1), (ic[3] - ic[1]) / (L - 1)
    w, h = abs(hx), abs(hy)
    xs, ys = ic[0] + np.arange(C) * hx, ic[1] + np.arange(L) * hy
    checked = 0
    for a in range(A):
        th = sc[0] + a * ((sc[2] - sc[0]) / (A - 1))
        c, s = math.cos(th), math.sin(th)
        if not _pixel_driven(c, s, hx, hy):
            continue
        p = 0.5 * (w * abs(c) + h * abs(s))
        q = 0.5 * abs(w * abs(c) - h * abs(s))
        lmax = min(w / abs(s), h / abs(c))
        for r in range(R):
            t = sc[1] + r * ((sc[3] - sc[1]) / (R - 1))
            ij, ln = O.dda_ray_path(shape, R, A, a, r, ic[:2], ic[2:], sc[:2], sc[2:])
            got = np.zeros(shape)
            np.add.at(got, (ij[:, 0], ij[:, 1]), ln)
            d = np.abs(t - (xs[None, :] * c + ys[:, None] * s))
            trap = lmax * np.clip((p - d) / (p - q), 0.0, 1.0)
            assert np.abs(got - trap).max() <= 1e-12, (a, r)
        checked += 1
    assert checked >= 3


@pytest.mark.parametrize("geom", GEOMS[:4], ids=lambda g: "%dx%d" % g[0])
def test_adjoint_and_normalisation(geom):
    shape, R, A, ic, sc = geom
    geo = _geo(ic, sc)
    rng = np.random.default_rng(5)
    x, y = rng.random(shape), rng.random((R, A))
    lhs = np.sum(O.radon_bresenham(x, R, A, **geo) * y)
    rhs = np.sum(x * O.radon_transpose_bresenham(y, shape, **geo))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    b = O.backprojection_bresenham(np.full((R, A), 2.5), shape, **geo)
    w = O.radon_transpose_bresenham(np.ones((R, A)), shape, **geo)
    assert np.allclose(b[w > 0], 2.5, rtol=1e-13) and np.all(np.isnan(b[w == 0]))


@pytest.fixture(scope="module")
def lib():
    from pyraft_amd.raftypes import libraft
    return libraft


def host_ray(lib, shape, R, A, a, r, ic, sc):
    cap = 2 * (shape[0] + shape[1]) + 8
    ij = np.zeros((cap, 2), np.int32)
    ln = np.zeros(cap)
    k = lib.raft_dda_host_ray(shape[0], shape[1], R, A, (ctypes.c_double * 4)(*ic), (ctypes.c_double * 4)(*sc), a, r,
                              ij.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                              ln.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), cap)
    assert k >= 0
    return ij[:k], ln[:k]


@pytest.mark.parametrize("geom", GEOMS + [((64, 64), 64, 181, DEF_IC, DEF_SC)], ids=lambda g: "%dx%d_A%d" % (g[0] + g[2:3]))
def test_device_traversal_matches_oracle(lib, geom):
    """The kernel's Radon per ray (host build) against the oracle, every ray: the same
    pixel sequence where the kernel traces the FP64 drift; on the exact-geometry angles,
    which the kernel sums in trapezoid form (two candidates per fast index, Q32 slow
    coordinate), the same lengths per pixel to 1e-6 of a full length (the Q32 coordinate
    is within ~N 2^-33 slow pixels of the exact one)."""
    shape, R, A, ic, sc = geom
    L, C = shape
    hx, hy = (ic[2] - ic[0]) / (C - 1), (ic[3] - ic[1]) / (L - 1)
    moved = 0
    for a in range(A):
        th = sc[0] + a * ((sc[2] - sc[0]) / (A - 1))
        fixed = _pixel_driven(math.cos(th), math.sin(th), hx, hy)
        for r in range(R):
            p1, l1 = host_ray(lib, shape, R, A, a, r, ic, sc)
            p2, l2 = O.dda_ray_path(shape, R, A, a, r, ic[:2], ic[2:], sc[:2], sc[2:])
            assert np.all((p1[:, 0] >= 0) & (p1[:, 0] < L) & (p1[:, 1] >= 0) & (p1[:, 1] < C))
            if not fixed:
                assert np.array_equal(p1, p2), (a, r)
                if len(l2):
                    assert np.abs(l1 - l2).max() <= 1e-13 * np.abs(l2).max(), (a, r)
                continue
            m1, m2 = np.zeros(shape), np.zeros(shape)
            np.add.at(m1, (p1[:, 0], p1[:, 1]), l1)
            np.add.at(m2, (p2[:, 0], p2[:, 1]), l2)
            full = max(abs(hx), abs(hy)) * 1.5
            assert np.abs(m1 - m2).max() <= 1e-6 * full, (a, r)
            moved += not np.array_equal(p1, p2)
    print("rays whose pixel sequence differs only by near-zero lengths:", moved, "of", R * A)


def test_host_ray_rejects_bad_geometry(lib):
    ij = np.zeros((4, 2), np.int32)
    ln = np.zeros(4)
    args = (ij.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ln.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 4)
    ic, sc = (ctypes.c_double * 4)(*DEF_IC), (ctypes.c_double * 4)(*DEF_SC)
    assert lib.raft_dda_host_ray(1, 8, 8, 8, ic, sc, 0, 0, *args) < 0
    assert lib.raft_dda_host_ray(8, 8, 8, 1, ic, sc, 0, 0, *args) < 0
    assert lib.raft_dda_host_ray(8, 8, 8, 8, ic, sc, 8, 0, *args) < 0


# ---------------------------------------------------------------- GPU


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


def _nan_match(got, ref):
    return np.array_equal(np.isnan(got), np.isnan(ref))


@pytest.mark.gpu
def test_gpu_golden(torch_mod, golden):
    torch = torch_mod
    from pyraft_amd.device import DdaPlan
    for c in DDA_CASES:
        print("dda golden", c, flush=True)
        shape, R, A, ic, sc = _golden_case(golden, c)
        p = DdaPlan(shape, R, A, ic, sc)
        img = torch.from_numpy(golden[c + "/img"].astype(np.float32)).cuda()
        sino = torch.from_numpy(golden[c + "/sino"].astype(np.float32)).cuda()
        assert rel_l2(p.radon(img).cpu().numpy(), golden[c + "/radon"]) <= REL_L2, c
        assert rel_l2(p.radon_transpose(sino).cpu().numpy(), golden[c + "/transpose"]) <= REL_L2, c
        bp = p.backprojection(sino).cpu().numpy()
        ref = golden[c + "/backprojection"]
        assert _nan_match(bp, ref), c
        ok = ~np.isnan(ref)
        assert rel_l2(bp[ok], ref[ok]) <= REL_L2, c


@pytest.mark.gpu
@pytest.mark.parametrize("shape,R,A,ic,sc", [
    ((128, 128), 150, 121, DEF_IC, DEF_SC),
    ((100, 77), 90, 60, (-0.7, 1.3, 1.2, -0.9), (-math.pi / 2, 1.1, math.pi / 2, -0.8)),
    ((256, 256), 256, 181, DEF_IC, (0.0, 1.0, math.pi, -1.0)),
    # rays narrower than the image: pixels beyond both edges of the ray range, and in
    # the second case pixels more than 64 rays beyond them
    ((96, 96), 60, 47, DEF_IC, (0.0, 0.5, math.pi, -0.5)),
    ((160, 160), 41, 37, DEF_IC, (0.0, 0.2, math.pi, -0.2)),
    # rays about 6 pixels apart: a wave's rays span more than one staged segment, so the
    # Radon reads the image directly
    ((256, 256), 40, 33, DEF_IC, DEF_SC),
])
def test_gpu_random_vs_oracle(torch_mod, shape, R, A, ic, sc):
    torch = torch_mod
    from pyraft_amd.device import DdaPlan
    print("dda random", shape, R, A, flush=True)
    geo = _geo(ic, sc)
    rng = np.random.default_rng(R + A)
    imgs = rng.random((3,) + shape).astype(np.float32)
    sinos = rng.random((3, R, A)).astype(np.float32)
    p = DdaPlan(shape, R, A, ic, sc)
    got_r = p.radon(torch.from_numpy(imgs).cuda()).cpu().numpy()
    got_t = p.radon_transpose(torch.from_numpy(sinos).cuda()).cpu().numpy()
    got_b = p.backprojection(torch.from_numpy(sinos).cuda()).cpu().numpy()
    for i in range(3):
        assert rel_l2(got_r[i], O.radon_bresenham(imgs[i].astype(np.float64), R, A, **geo)) <= REL_L2, i
        assert rel_l2(got_t[i], O.radon_transpose_bresenham(sinos[i].astype(np.float64), shape, **geo)) <= REL_L2, i
        ref = O.backprojection_bresenham(sinos[i].astype(np.float64), shape, **geo)
        assert _nan_match(got_b[i], ref), i
        ok = ~np.isnan(ref)
        assert rel_l2(got_b[i][ok], ref[ok]) <= REL_L2, i
    w = p.weights().cpu().numpy()
    assert rel_l2(w, O.radon_transpose_bresenham(np.ones((R, A)), shape, **geo)) <= REL_L2


@pytest.mark.gpu
def test_gpu_transpose_tall_tiles_vs_oracle(torch_mod):
    """A launch of at least 256 gather blocks takes the 64 x 64-pixel tiles (16 lines
    per thread): a 1024^2 image against the oracle's transpose and backprojection."""
    torch = torch_mod
    from pyraft_amd.device import DdaPlan
    n, R, A = 1024, 1024, 48
    geo = _geo(DEF_IC, DEF_SC)
    y = np.random.default_rng(12).random((R, A)).astype(np.float32)
    p = DdaPlan((n, n), R, A)
    got_t = p.radon_transpose(torch.from_numpy(y).cuda()).cpu().numpy()
    got_b = p.backprojection(torch.from_numpy(y).cuda()).cpu().numpy()
    y64 = y.astype(np.float64)
    assert rel_l2(got_t, O.radon_transpose_bresenham(y64, (n, n), **geo)) <= REL_L2
    ref = O.backprojection_bresenham(y64, (n, n), **geo)
    assert _nan_match(got_b, ref)
    ok = ~np.isnan(ref)
    assert rel_l2(got_b[ok], ref[ok]) <= REL_L2


@pytest.mark.gpu
def test_gpu_full_size_vs_oracle(torch_mod):
    """2048^2 image, 2048 rays (the benchmark's ray spacing, so the same staged segments
    and 8-row chunks over the full fast range), 181 angles: Radon and transpose against
    the oracle."""
    torch = torch_mod
    from pyraft_amd.device import DdaPlan
    shape, R, A = (2048, 2048), 2048, 181
    print("dda full size", shape, R, A, flush=True)
    geo = _geo(DEF_IC, DEF_SC)
    rng = np.random.default_rng(11)
    img = rng.random((1,) + shape).astype(np.float32)
    sino = rng.random((1, R, A)).astype(np.float32)
    p = DdaPlan(shape, R, A, DEF_IC, DEF_SC)
    got_r = p.radon(torch.from_numpy(img).cuda()).cpu().numpy()
    got_t = p.radon_transpose(torch.from_numpy(sino).cuda()).cpu().numpy()
    assert rel_l2(got_r[0], O.radon_bresenham(img[0].astype(np.float64), R, A, **geo)) <= REL_L2
    assert rel_l2(got_t[0], O.radon_transpose_bresenham(sino[0].astype(np.float64), shape, **geo)) <= REL_L2


@pytest.mark.gpu
def test_gpu_straight_angles(torch_mod):
    """Square pixels, 181 angles over [0, pi]: the axis angles 0, pi/2 and pi run straight
    (one column or line per ray, full length), so the transpose sums them per column or
    line instead of tracing them; with non-square pixels the traced angles bend and stay
    traced."""
    from pyraft_amd.device import DdaPlan
    info = DdaPlan((256, 256), 256, 181).info()
    assert info["straight_angles"] == 3 and info["traced_angles"] == 0, info
    info = DdaPlan((100, 77), 90, 60, (-0.7, 1.3, 1.2, -0.9), (-math.pi / 2, 1.1, math.pi / 2, -0.8)).info()
    assert info["straight_angles"] == 0 and info["traced_angles"] > 0, info


@pytest.mark.gpu
@pytest.mark.parametrize("shape,R,A", [((256, 256), 256, 181), ((2048, 2048), 2048, 1500)])
def test_gpu_adjoint(torch_mod, shape, R, A):
    torch = torch_mod
    from pyraft_amd.device import DdaPlan
    print("dda adjoint", shape, R, A, flush=True)
    p = DdaPlan(shape, R, A)
    g = torch.Generator("cuda").manual_seed(3)
    x = torch.rand(shape, device="cuda", generator=g)
    y = torch.rand((R, A), device="cuda", generator=g)
    lhs = float((p.radon(x).double() * y.double()).sum())
    rhs = float((x.double() * p.radon_transpose(y).double()).sum())
    print("adjoint mismatch", abs(lhs - rhs) / abs(lhs), flush=True)
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
    # Radon of a constant image: the chord through the pixel-edge box (square pixels)
    ones = p.radon(torch.ones(shape, device="cuda")).cpu().numpy().astype(np.float64)
    h = 2.0 / (shape[0] - 1)
    ths = np.arange(A) * (math.pi / (A - 1))
    ts = 1.0 + np.arange(R) * (-2.0 / (R - 1))
    pick = np.random.default_rng(1).integers(0, [R, A], size=(400, 2))
    for r, a in pick:
        ref = clip_len(math.cos(ths[a]), math.sin(ths[a]), ts[r], -1 - h / 2, 1 + h / 2, -1 - h / 2, 1 + h / 2)
        assert abs(ones[r, a] - ref) <= 1e-6 * max(ref, 1.0), (r, a)


@pytest.mark.gpu
def test_gpu_dropins_and_wrappers(torch_mod):
    import pyraft_amd as pr
    from pyraft_amd.raftypes import check_legacy, libraft, make_RAFT_IMAGE
    n, R, A = 64, 70, 45
    rng = np.random.default_rng(9)
    f = pr.image(rng.random((n, n)))
    s = pr.radon(f, (R, A))                         # pyraft's default method
    geo = _geo(DEF_IC, DEF_SC)
    assert rel_l2(s, O.radon_bresenham(np.asarray(f), R, A, **geo)) <= REL_L2
    y = pr.image(rng.random((R, A)), x_extent=(0.0, math.pi))
    g = pr.radon_transpose(y, (n, n))
    assert rel_l2(g, O.radon_transpose_bresenham(np.asarray(y), (n, n), **geo)) <= REL_L2
    b = pr.backprojection(y, pr.backprojection_plan(y.shape, (n, n), method=pr.radon_method.BRESENHAM))
    ref = O.backprojection_bresenham(np.asarray(y), (n, n), **geo)
    assert _nan_match(np.asarray(b), ref) and rel_l2(np.asarray(b)[~np.isnan(ref)], ref[~np.isnan(ref)]) <= REL_L2
    # the drop-in transpose adds into the caller's image; the backprojection divides the sum
    base = rng.random((n, n))
    img = base.copy()
    libraft.raft_radon_transpose_bresenham(make_RAFT_IMAGE(np.asarray(y), (0.0, 1.0), (math.pi, -1.0)),
                                           make_RAFT_IMAGE(img), 4)
    check_legacy("raft_radon_transpose_bresenham")
    assert rel_l2(img - base, O.radon_transpose_bresenham(np.asarray(y), (n, n), **geo)) <= REL_L2
    img = base.copy()
    libraft.raft_backprojection_bresenham(make_RAFT_IMAGE(np.asarray(y), (0.0, 1.0), (math.pi, -1.0)),
                                          make_RAFT_IMAGE(img), 4)
    check_legacy("raft_backprojection_bresenham")
    w = O.radon_transpose_bresenham(np.ones((R, A)), (n, n), **geo)
    with np.errstate(all="ignore"):
        ref = (base + O.radon_transpose_bresenham(np.asarray(y), (n, n), **geo)) / w
    assert _nan_match(img, ref) and rel_l2(img[~np.isnan(ref)], ref[~np.isnan(ref)]) <= REL_L2


@pytest.mark.gpu
def test_gpu_em_default_method(torch_mod):
    """pyraft.em with its defaults: x <- x R^T(s / R x) / R^T 1 on the line-traversal pair."""
    import pyraft_amd as pr
    n, R, A = 48, 60, 40
    rng = np.random.default_rng(4)
    truth = rng.random((n, n)) + 0.5
    geo = _geo(DEF_IC, DEF_SC)
    s = pr.image(O.radon_bresenham(truth, R, A, **geo), x_extent=(0.0, math.pi))
    x = pr.em(s, (n, n), niter=3)
    ref = np.ones((n, n))
    w = O.radon_transpose_bresenham(np.ones((R, A)), (n, n), **geo)
    with np.errstate(all="ignore"):
        for _ in range(3):
            ref = ref * O.radon_transpose_bresenham(np.asarray(s) / O.radon_bresenham(ref, R, A, **geo), (n, n),
                                                    **geo) / w
    ok = np.isfinite(ref)
    assert np.array_equal(ok, np.isfinite(np.asarray(x)))
    assert rel_l2(np.asarray(x)[ok], ref[ok]) <= REL_L2
