"""Geometry sweep of both projector pairs on the GPU against the FP64 oracle.

Each slant-stack geometry exercises one departure from the default field of view
(raft/raft_radon_slantstack.c:44-214 reads every pitch with its sign and every range
from the corners): flipped image axes (dx < 0, dy > 0), decreasing angles (dth < 0),
a decreasing detector coordinate (dt < 0), 4:1 pixels, angle ranges beyond one
half-turn, detectors off the image centre (rays that miss the image), few wide bins
and many narrow ones, and an image away from the origin.  For each: Radon,
backprojection and the exact transpose within rel-L2 1e-5 of the oracle, the adjoint
identity to 1e-6, the FBP ramp filter of that detector range, and two EM iterations.
All geometries then run as one uneven batch, each slice against the oracle.

The line-traversal pair runs the CPU suite's geometries (flipped axes, rectangular
images, rays off the image, axis-only angle sets) on the device.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5
ADJOINT = 1e-6
PI = math.pi
DEF_IC = (-1.0, 1.0, 1.0, -1.0)
DEF_SC = (0.0, 1.0, PI, -1.0)

# n, R, A, image corners (tl_x, tl_y, br_x, br_y), sinogram corners
SWEEP = [
    (37, 41, 29, (1.0, 1.0, -1.0, -1.0), DEF_SC),                  # x axis flipped
    (50, 33, 60, (-1.0, -1.0, 1.0, 1.0), DEF_SC),                  # y axis flipped
    (44, 90, 17, (1.2, -0.8, -0.9, 1.1), (-PI, 0.9, 0.3, -1.3)),   # both flipped, detector off-centre
    (29, 64, 50, (-2.0, 0.5, 2.0, -0.5), (0.0, 2.1, PI, -2.1)),    # 4:1 pixels
    (70, 25, 111, DEF_IC, (PI, 1.0, 0.0, -1.0)),                   # decreasing angles
    (33, 40, 45, DEF_IC, (0.0, -1.0, PI, 1.0)),                    # decreasing detector coordinate
    (64, 128, 97, DEF_IC, (-PI, 1.0, 2.0 * PI, -1.0)),             # three half-turns
    (80, 60, 36, DEF_IC, (0.0, 0.3, PI, -1.7)),                    # detector off-centre: rays miss
    (100, 7, 40, DEF_IC, (0.0, 3.0, PI, -3.0)),                    # few wide bins
    (25, 120, 12, DEF_IC, (0.0, 0.4, PI, -0.4)),                   # many narrow bins
    (96, 96, 64, (-0.3, 5.2, 0.9, 4.0), (0.25, 5.5, 3.0, 3.7)),    # image away from the origin
]


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def _geo(ic, sc):
    return dict(img_tl=(ic[0], ic[1]), img_br=(ic[2], ic[3]), sino_tl=(sc[0], sc[1]), sino_br=(sc[2], sc[3]))


def _ids(g):
    return "n%d_R%d_A%d" % g[:3]


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.mark.parametrize("geom", SWEEP, ids=_ids)
def test_slantstack_geometry(torch_mod, geom):
    torch = torch_mod
    from pyraft_amd.device import Plan
    n, R, A, ic, sc = geom
    geo = _geo(ic, sc)
    print("slant-stack sweep", geom, flush=True)
    rng = np.random.default_rng(n * 1000 + R * 10 + A)
    x = rng.random((n, n)).astype(np.float32)
    y = rng.random((R, A)).astype(np.float32)
    p = Plan(n, R, A, ic, sc)
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    Rx = p.radon(xd).cpu().numpy()
    e_r = rel_l2(Rx, O.radon(x.astype(np.float64), R, A, **geo))
    e_b = rel_l2(p.backprojection(yd).cpu().numpy(), O.backprojection(y.astype(np.float64), n, **geo))
    Rty = p.radon_transpose(yd).cpu().numpy()
    e_t = rel_l2(Rty, O.radon_transpose(y.astype(np.float64), n, **geo))
    lhs = float(np.sum(Rx.astype(np.float64) * y))
    rhs = float(np.sum(x.astype(np.float64) * Rty))
    e_f = rel_l2(p.fbp_filter(yd).cpu().numpy(), O.lowpassino_fbp(y.astype(np.float64), geo["sino_tl"],
                                                                  geo["sino_br"]))
    print("radon", e_r, "backprojection", e_b, "transpose", e_t, "adjoint", abs(lhs - rhs) / max(abs(lhs), 1e-30),
          "filter", e_f, flush=True)
    assert e_r <= REL_L2 and e_b <= REL_L2 and e_t <= REL_L2 and e_f <= REL_L2
    assert abs(lhs - rhs) <= ADJOINT * max(abs(lhs), 1e-30)
    s = (Rx + 0.5).astype(np.float32)  # positive data
    got = p.em(torch.from_numpy(s).cuda(), 2).cpu().numpy()
    ref, _ = O.osem(s.astype(np.float64), n, 2, **geo)
    e_em = rel_l2(got, ref)
    print("em", e_em, flush=True)
    assert e_em <= REL_L2


def test_slantstack_sweep_as_one_batch(torch_mod):
    torch = torch_mod
    from pyraft_amd.batch import Batch
    b = Batch([g[0] for g in SWEEP], [g[1] for g in SWEEP], [g[2] for g in SWEEP],
              img_corners=[g[3] for g in SWEEP], sino_corners=[g[4] for g in SWEEP])
    rng = np.random.default_rng(77)
    imgs = [rng.random((g[0], g[0])).astype(np.float32) for g in SWEEP]
    sinos = [rng.random((g[1], g[2])).astype(np.float32) for g in SWEEP]
    rad = b.radon(b.pack_images([torch.from_numpy(v).cuda() for v in imgs]))
    bp = b.backprojection(b.pack_sinos([torch.from_numpy(v).cuda() for v in sinos]))
    for i, (n, R, A, ic, sc) in enumerate(SWEEP):
        geo = _geo(ic, sc)
        e_r = rel_l2(b.sino(rad, i).cpu().numpy(), O.radon(imgs[i].astype(np.float64), R, A, **geo))
        e_b = rel_l2(b.image(bp, i).cpu().numpy(), O.backprojection(sinos[i].astype(np.float64), n, **geo))
        print("batch slice", i, SWEEP[i][:3], "radon", e_r, "backprojection", e_b, flush=True)
        assert e_r <= REL_L2 and e_b <= REL_L2, i


# (lines, cols), R, A, image corners, sinogram corners
DDA_SWEEP = [
    ((16, 16), 12, 9, DEF_IC, (0.0, 1.0, PI, -1.0)),
    ((33, 33), 40, 31, DEF_IC, (0.0, 1.0, PI, -1.0)),
    ((20, 31), 25, 17, (-0.7, 1.3, 1.2, -0.9), (-PI / 2, 1.1, PI / 2, -0.8)),   # non-square pixels
    ((24, 24), 30, 45, (1.0, -1.0, -1.0, 1.0), (PI, -1.2, 0.0, 1.2)),           # flipped axes
    ((40, 40), 41, 4, DEF_IC, (0.0, 1.0, PI / 2, -1.0)),                       # 0, 30, 60, 90 degrees
    ((17, 29), 19, 13, DEF_IC, (0.0, 1.5, 2 * PI, -1.5)),                      # rays off the image
    ((64, 20), 70, 33, (-0.5, 2.0, 0.5, -2.0), (0.1, 2.2, 3.0, -2.2)),          # tall image
]


def _nan_match(a, b):
    return np.array_equal(np.isnan(a), np.isnan(b))


@pytest.mark.parametrize("geom", DDA_SWEEP, ids=lambda g: "%dx%d_R%d_A%d" % (g[0] + g[1:3]))
def test_line_traversal_geometry(torch_mod, geom):
    torch = torch_mod
    from pyraft_amd.device import DdaPlan
    shape, R, A, ic, sc = geom
    geo = _geo(ic, sc)
    print("line-traversal sweep", geom, flush=True)
    rng = np.random.default_rng(shape[0] * shape[1] + R * A)
    img = rng.random(shape).astype(np.float32)
    sino = rng.random((R, A)).astype(np.float32)
    p = DdaPlan(shape, R, A, ic, sc)
    e_r = rel_l2(p.radon(torch.from_numpy(img).cuda()).cpu().numpy(),
                 O.radon_bresenham(img.astype(np.float64), R, A, **geo))
    e_t = rel_l2(p.radon_transpose(torch.from_numpy(sino).cuda()).cpu().numpy(),
                 O.radon_transpose_bresenham(sino.astype(np.float64), shape, **geo))
    bp = p.backprojection(torch.from_numpy(sino).cuda()).cpu().numpy()
    ref = O.backprojection_bresenham(sino.astype(np.float64), shape, **geo)
    ok = ~np.isnan(ref)
    e_b = rel_l2(bp[ok], ref[ok])
    print("radon", e_r, "transpose", e_t, "backprojection", e_b, flush=True)
    assert e_r <= REL_L2 and e_t <= REL_L2
    assert _nan_match(bp, ref) and e_b <= REL_L2
