"""Edge sizes of the HIP slant-stack path: the smallest geometries (n = 2, R = 2, A = 1,
...) against the FP64 oracle, and the largest supported image side (n = 16384) through
size-independent closed forms.

At n = R = 16384 with the default corners the rays of theta = 0 sit on the pixel
columns (t_k = x_k, slope 0, one cell per ray), so ray k is the sum of column k times
dx; at theta = pi/2 it is the sum of row k times dy.  The backprojection of a two-angle
sinogram is then (s[k][0] + s[j][1]) * pi/2 at pixel (j, k).  Columns and rows at the
border are left out: there the FP64 predicate's rounding decides which of two taps
carries the sample.  The transpose is checked by the adjoint identity.
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


SMALL = [(2, 2, 1), (3, 2, 2), (2, 5, 3), (4, 3, 7), (5, 9, 4), (7, 4, 13)]


@pytest.mark.parametrize("geom", SMALL, ids=lambda g: "n%d_R%d_A%d" % g)
def test_smallest_geometries(torch_mod, geom):
    torch = torch_mod
    from pyraft_amd.device import Plan
    n, R, A = geom
    print("small geometry", geom, flush=True)
    rng = np.random.default_rng(n * 100 + R * 10 + A)
    x = rng.random((n, n)).astype(np.float32)
    y = rng.random((R, A)).astype(np.float32)
    p = Plan(n, R, A)
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    Rx = p.radon(xd).cpu().numpy()
    assert rel_l2(Rx, O.radon(x.astype(np.float64), R, A)) <= REL_L2
    assert rel_l2(p.backprojection(yd).cpu().numpy(), O.backprojection(y.astype(np.float64), n)) <= REL_L2
    Rty = p.radon_transpose(yd).cpu().numpy()
    assert rel_l2(Rty, O.radon_transpose(y.astype(np.float64), n)) <= REL_L2
    lhs = float(np.sum(Rx.astype(np.float64) * y))
    rhs = float(np.sum(x.astype(np.float64) * Rty))
    assert abs(lhs - rhs) <= ADJOINT * max(abs(lhs), 1e-30)
    assert rel_l2(p.fbp_filter(yd).cpu().numpy(), O.lowpassino_fbp(y.astype(np.float64))) <= REL_L2
    s = (Rx + 0.5).astype(np.float32)  # positive data
    got = p.em(torch.from_numpy(s).cuda(), 2).cpu().numpy()
    ref, _ = O.osem(s.astype(np.float64), n, 2)
    assert rel_l2(got, ref) <= REL_L2


def test_largest_image_side(torch_mod):
    torch = torch_mod
    from pyraft_amd.device import Plan
    n = R = 16384
    A = 2
    print("largest image side", n, flush=True)
    gen = torch.Generator(device="cuda").manual_seed(16384)
    x = torch.rand((n, n), device="cuda", generator=gen)
    p = Plan(n, R, A)
    sino = p.radon(x)
    xs = x.double()
    col = (xs.sum(dim=0) * (2.0 / (n - 1))).cpu().numpy()   # column sums * dx
    row = (xs.sum(dim=1) * (2.0 / (n - 1))).cpu().numpy()   # row sums * dy
    s = sino.cpu().numpy().astype(np.float64)
    inner = slice(1, n - 1)
    print("theta 0", rel_l2(s[inner, 0], col[inner]), "theta pi/2", rel_l2(s[inner, 1], row[inner]), flush=True)
    assert rel_l2(s[inner, 0], col[inner]) <= REL_L2
    assert rel_l2(s[inner, 1], row[inner]) <= REL_L2
    # backprojection of the two-angle sinogram: (s[k][0] + s[j][1]) * pi / 2 at (j, k)
    bp = p.backprojection(sino)
    sd = sino.double()
    expect = (sd[:, 0][None, :] + sd[:, 1][:, None]) * (math.pi / 2)
    d = (bp.double() - expect)[1:n - 1, 1:n - 1]
    e = float(torch.linalg.vector_norm(d) / torch.linalg.vector_norm(expect[1:n - 1, 1:n - 1]))
    print("backprojection", e, flush=True)
    assert e <= REL_L2
    del bp, expect, d
    y = torch.rand((R, A), device="cuda", generator=gen)
    lhs = float((sd * y.double()).sum())
    rhs = float((xs * p.radon_transpose(y).double()).sum())
    print("adjoint", abs(lhs - rhs) / abs(lhs), flush=True)
    assert abs(lhs - rhs) <= ADJOINT * abs(lhs)
