"""Pins for the FP64 oracle (CPU only).

1. Golden vectors: outputs of the reference's own C (tests/golden/make_golden.py)
   must be reproduced bit for bit.
2. The reference checker oracle/_ref, when built here, on fresh random inputs.
3. Closed forms / invariants that a plausible mistake would break:
   row and column sums at theta = pi/2 and theta = 0 (catches transposed image
   axes and wrong scale), the knife-edge exclusion of the last detector bin, chord
   lengths of a constant image, BP(1) = pi inside the unit disk, the adjoint
   identity of the transpose, the analytic Radon of an ellipse (orientation), the
   filter's impulse response and DC gain, and the EM fixed point.
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "slantstack_golden.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


def _cases(golden):
    """slant-stack cases (the line-traversal ones, dda_*, are checked in test_bresenham.py)"""
    return sorted({k.split("/")[0] for k in golden.files if k.endswith("/geom") and not k.startswith("dda_")})


def _geo(g):
    n, R, A = int(g[0]), int(g[1]), int(g[2])
    ic, sc = tuple(g[3:7]), tuple(g[7:11])
    return n, R, A, dict(img_tl=ic[:2], img_br=ic[2:], sino_tl=sc[:2], sino_br=sc[2:])


def test_golden_radon_backprojection(golden):
    cases = _cases(golden)
    assert len(cases) >= 6
    for c in cases:
        n, R, A, geo = _geo(golden[c + "/geom"])
        r = O.radon(golden[c + "/img"], R, A, nthreads=3, **geo)
        assert np.array_equal(r, golden[c + "/radon"]), c
        b = O.backprojection(golden[c + "/sino"], n, nthreads=3, **geo)
        assert np.array_equal(b, golden[c + "/backprojection"]), c


def test_golden_phantoms(golden):
    d = golden["shepp/desc"]
    assert np.array_equal(O.phantom_fromdesc(d, 128), golden["shepp/phantom128"])
    assert np.array_equal(O.phantom_fromdesc(d, 33, tl=(-0.9, 1.1), br=(1.2, -1.0)), golden["shepp/phantom33_fov"])
    assert np.array_equal(O.radon_fromdesc(d, 96, 60), golden["shepp/radon_fromdesc_96x60"])
    assert np.array_equal(O.radon_fromdesc(d, 40, 7, tl=(0.2, 1.2), br=(2.9, -0.7)),
                          golden["shepp/radon_fromdesc_40x7_fov"])


@pytest.mark.skipif(not O.ref_available(), reason="reference checker oracle/_ref not built here")
@pytest.mark.parametrize("n,R,A", [(31, 31, 90), (40, 52, 33), (128, 128, 60)])
def test_against_reference_c(n, R, A):
    rng = np.random.default_rng(n * 1000 + A)
    img = rng.random((n, n))
    sino = rng.random((R, A))
    assert np.array_equal(O.radon(img, R, A), O.ref_radon(img, R, A, nthreads=4))
    assert np.array_equal(O.backprojection(sino, n), O.ref_backprojection(sino, n, nthreads=4))


def test_row_and_column_sums():
    """theta = 0: t_k on the column grid -> s[k,0] = dx * sum_i f[i,k] (k <= N-2);
    theta = pi/2: t_k on the row grid  -> s[k,A/2] = dy * sum_i f[k,i] (k <= N-2);
    the t = +1 ray lands exactly on u = N-1 and is excluded at theta = 0."""
    n = R = 32
    A = 8
    rng = np.random.default_rng(5)
    f = rng.random((n, n))
    s = O.radon(f, R, A)
    d = 2.0 / (n - 1)
    np.testing.assert_allclose(s[:n - 1, 0], d * f[:, :n - 1].sum(axis=0), rtol=1e-12)
    assert s[n - 1, 0] == 0.0
    np.testing.assert_allclose(s[:n - 1, A // 2], d * f[:n - 1, :].sum(axis=1), rtol=1e-9)


def test_constant_image_chords():
    n = R = 401
    A = 12
    s = O.radon(np.ones((n, n)), R, A)
    k0 = (R - 1) // 2  # t = 0
    for j in range(A):
        th = j * math.pi / A
        chord = 2.0 / max(abs(math.cos(th)), abs(math.sin(th)))
        assert abs(s[k0, j] - chord) <= 2 * (2.0 / (n - 1)) / min(1, max(abs(math.cos(th)), abs(math.sin(th)))), j


def test_bp_of_ones_is_pi_inside_unit_disk():
    n = R = 64
    A = 90
    b = O.backprojection(np.ones((R, A)), n)
    d = 2.0 / (n - 1)
    y, x = np.meshgrid(-1 + d * np.arange(n), -1 + d * np.arange(n), indexing="ij")
    inside = np.sqrt(x * x + y * y) < 1 - 2 * d
    np.testing.assert_allclose(b[inside], math.pi, rtol=1e-12)
    assert np.all(b <= math.pi + 1e-12)


def test_single_angle_backprojection_is_a_ridge():
    """One non-zero column j: BP = dtheta * s_j interpolated at t = x cos + y sin."""
    n = R = 40
    A = 5
    j = 2
    s = np.zeros((R, A))
    col = np.random.default_rng(1).random(R)
    s[:, j] = col
    b = O.backprojection(s, n)
    dth = math.pi / A
    th = j * dth
    d = 2.0 / (n - 1)
    for r in (3, 17, 30):
        for c in (5, 20, 33):
            t = (-1 + c * d) * math.cos(th) + (-1 + r * d) * math.sin(th)
            expect = dth * np.interp(t, -1 + d * np.arange(R), col) if -1 <= t < 1 else 0.0
            assert abs(b[r, c] - expect) < 1e-12


@pytest.mark.parametrize("n,R,A", [(40, 40, 30), (33, 47, 19)])
def test_transpose_adjoint_identity(n, R, A):
    rng = np.random.default_rng(R)
    x = rng.random((n, n))
    y = rng.random((R, A))
    lhs = np.sum(O.radon(x, R, A) * y)
    rhs = np.sum(x * O.radon_transpose(y, n))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_radon_matches_analytic_ellipses():
    """The slant-stack puts image row 0 at y = ymin, the rasteriser at y = +1: the
    slant-stack of the raster is the Radon of the vertically flipped description,
    sampled on the theta = j*pi/A, ascending-t grid."""
    n = R = 256
    A = 30
    d = np.array([[1.0, 0.5, 0.3, 0.2, 0.3, 0.4], [0.6, 0.2, 0.25, -0.3, -0.2, 1.2]])
    s = O.radon(O.phantom_fromdesc(d, n), R, A)
    flip = d.copy()
    flip[:, 4] *= -1
    flip[:, 5] *= -1
    exact = O.radon_fromdesc(flip, R, A, tl=(0.0, -1.0), br=(math.pi * (1 - 1.0 / A), 1.0))
    err = np.linalg.norm(s - exact) / np.linalg.norm(exact)
    assert err < 2e-2, err
    # and the unflipped description does not match (orientation is pinned)
    wrong = O.radon_fromdesc(d, R, A, tl=(0.0, -1.0), br=(math.pi * (1 - 1.0 / A), 1.0))
    assert np.linalg.norm(s - wrong) / np.linalg.norm(wrong) > 0.2


@pytest.mark.parametrize("R", [64, 65])
def test_filter_impulse_response_and_dc_gain(R):
    """lowpassino_fbp is circular convolution with g = Re(ifft(fftshift(h))): a unit
    impulse at row 0 returns g, a constant returns G[0] times the constant."""
    imp = np.zeros((R, 1))
    imp[0, 0] = 1.0
    g = O.lowpassino_fbp(imp)[:, 0]
    np.testing.assert_allclose(g, O.ramp_kernel(R), atol=1e-14)
    # shift invariance (circular): impulse at row 5 -> g rolled by 5
    imp5 = np.zeros((R, 1))
    imp5[5, 0] = 1.0
    np.testing.assert_allclose(O.lowpassino_fbp(imp5)[:, 0], np.roll(g, 5), atol=1e-12)
    dt = 2.0 / (R - 1)
    wc = 1 / (2 * dt)
    w = np.linspace(-wc, wc, R)
    G0 = np.fft.fftshift(np.abs(w) * np.cos(np.pi * w * dt))[0]
    np.testing.assert_allclose(O.lowpassino_fbp(np.full((R, 2), 3.0)), 3.0 * G0, rtol=1e-9)
    # g is real and even: g[k] == g[R-k]
    np.testing.assert_allclose(g[1:], g[1:][::-1], atol=1e-14)


def test_fbp_reconstructs_phantom_roughly():
    n = R = 128
    A = 180
    d = np.array([[1.0, 0.45, 0.35, 0.1, -0.1, 0.3]])
    f = O.phantom_fromdesc(d, n)
    rec = O.fbp(O.radon(f, R, A), n)
    # FBP scale: the slant-stack pair with this ramp reproduces f up to a constant
    c = np.sum(rec * f) / np.sum(f * f)
    err = np.linalg.norm(rec - c * f) / np.linalg.norm(c * f)
    assert err < 0.35 and c > 0, (err, c)


def test_em_fixed_point_and_positivity():
    n = R = 24
    A = 18
    rng = np.random.default_rng(9)
    xs = 0.5 + rng.random((n, n))
    s = O.radon(xs, R, A)
    x, ll = O.em(s, n, 3, x0=xs)
    np.testing.assert_allclose(x, xs, rtol=1e-12)
    x1, _ = O.osem(s, n, 2, nsubsets=3)
    assert np.all(x1 >= 0) and np.all(np.isfinite(x1))
    x0, ll0 = O.em(s, n, 0)
    assert np.array_equal(x0, np.ones((n, n))) and ll0 == []


@pytest.mark.parametrize("S", [3, 5, 16])
def test_osem_subsets_fixed_point(S):
    """OSEM with S contiguous subsets (pyraft/iterative/em.py:282-318 with the
    raft/raft_em_gpu.cu:63-69 update): with s = R x*, every subset update maps x* to
    itself only if subset b divides by its own weight BP_b(1) and projects exactly its
    own angle block.  Dividing by the full-range BP(1), or projecting the wrong block,
    moves x* by O(1)."""
    n = R = 24
    A = 48
    rng = np.random.default_rng(90 + S)
    xs = 0.5 + rng.random((n, n))
    s = O.radon(xs, R, A)
    x, ll = O.osem(s, n, 2, nsubsets=S, x0=xs)
    np.testing.assert_allclose(x, xs, rtol=1e-12)
    # the log-likelihood terms of the subsets add up to the full one: sum s log s - s
    m = s > 0
    full = float(np.sum(s[m] * np.log(s[m])) - np.sum(s[m]))
    assert len(ll) == 2 and all(abs(v - full) <= 1e-9 * abs(full) for v in ll), (ll, full)
    # the subset blocks tile the angles, and their weights add up to BP(1)
    blocks = O.subset_bounds(A, S)
    assert blocks[0][0] == 0 and blocks[-1][1] == A
    assert all(b[1] == c[0] for b, c in zip(blocks, blocks[1:]))
    ones = np.ones((R, A))
    wsum = sum(O.backprojection(ones, n, j0=b0, j1=b1) for b0, b1 in blocks)
    np.testing.assert_allclose(wsum, O.backprojection(ones, n), rtol=1e-12, atol=1e-12)
    # one iteration from a flat start moves toward x* and differs from plain EM
    x1, _ = O.osem(s, n, 1, nsubsets=S)
    e1, _ = O.em(s, n, 1)
    assert np.linalg.norm(x1 - xs) < np.linalg.norm(e1 - xs)


def test_osem_planted_weight_mistake_is_caught():
    """The check above fails for a plausible mistake planted in a copy of the
    recurrence: dividing every subset by the full-range weight."""
    n = R = 16
    A = 24
    S = 4
    rng = np.random.default_rng(7)
    xs = 0.5 + rng.random((n, n))
    s = O.radon(xs, R, A)
    W = O.backprojection(np.ones((R, A)), n)
    x = xs.copy()
    for b0, b1 in O.subset_bounds(A, S):
        p = O.radon(x, R, A, j0=b0, j1=b1)[:, b0:b1]
        rho = np.zeros((R, A))
        rho[:, b0:b1] = np.where(p > 0, s[:, b0:b1] / np.where(p > 0, p, 1.0), 0.0)
        x = np.where(W > 0, x * O.backprojection(rho, n, j0=b0, j1=b1) / np.where(W > 0, W, 1.0), x)
    assert np.linalg.norm(x - xs) > 0.5 * np.linalg.norm(xs)


def test_loglik_formula():
    s = np.array([[1.0, 2.0], [0.0, 3.0]])
    p = np.array([[2.0, 0.0], [1.0, 4.0]])
    expect = 1.0 * math.log(2.0) - 2.0 + 0.0 - 1.0 + 3.0 * math.log(4.0) - 4.0
    assert abs(O.loglik(s, p) - expect) < 1e-12


def test_backprojection_rows_equals_full_image():
    """oracle.backprojection_rows computes each listed row exactly as the full image does,
    including an angle sub-range and non-default corners."""
    rng = np.random.default_rng(12)
    n, R, A = 37, 41, 29
    sino = rng.random((R, A))
    geo = dict(img_tl=(-0.8, 1.1), img_br=(1.2, -0.9), sino_tl=(0.1, 1.2), sino_br=(3.0, -1.1))
    rows = [0, 1, 17, 35, 36, 17]
    full = O.backprojection(sino, n, **geo)
    assert np.array_equal(O.backprojection_rows(sino, n, rows, **geo), full[rows])
    part = O.backprojection(sino, n, j0=5, j1=20, **geo)
    assert np.array_equal(O.backprojection_rows(sino, n, rows, j0=5, j1=20, **geo), part[rows])
    with pytest.raises(ValueError):
        O.backprojection_rows(sino, n, [n], **geo)
