"""The 256-slice uneven batch (n from {256, ..., 2048}, A ~ U[180, 1800], R = n,
32-ellipse phantoms, seed 4) through the batched persistent kernels, checked against
the FP64 oracle on a stratified sample of its slices.

The reference projects one slice per call and loops over OSEM subsets in Python
(pyraft/sinogram/sino.py:284-311, pyraft/iterative/em.py:265-320); here all 256
slices run in one batched launch per kernel, captured into a HIP graph.  For every
image side the sample takes the slice with the fewest and the one with the most
angles (12 slices):

* Radon: every angle for n <= 768; for larger n, blocks of angles at theta = 0, around
  pi/4 (the branch switch), pi/2 and the last angles;
* backprojection over all angles: every pixel for n <= 512, a sample of rows (the
  edges, the centre and a stride) for larger n;
* OSEM, 8 contiguous subsets, 2 iterations on s = R f: against the oracle's recurrence
  for n <= 768 (log-likelihood included); for larger n against per-slice OSEM calls,
  whose kernels the oracle checks at 1024^2 / 720 angles / 16 subsets elsewhere.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
REL_L2 = 1e-5
SIZES = [256, 512, 768, 1024, 1536, 2048]


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def c4_geometry(nslices=256, seed=4):
    rng = np.random.default_rng(seed)
    sizes = [int(v) for v in rng.choice(SIZES, nslices)]
    angles = [int(v) for v in rng.integers(180, 1801, nslices)]
    return sizes, angles


def stratified(sizes, angles):
    """Per image side, the slice with the fewest and the one with the most angles."""
    pick = []
    for n in SIZES:
        idx = [i for i, v in enumerate(sizes) if v == n]
        assert idx, n
        lo = min(idx, key=lambda i: (angles[i], i))
        hi = max(idx, key=lambda i: (angles[i], -i))
        pick += [lo] if lo == hi else [lo, hi]
    return pick


@pytest.fixture(scope="module")
def c4():
    import torch
    from pyraft_amd.batch import Batch
    from pyraft_amd.device import phantom
    from pyraft_amd.phantom import random_ellipses
    assert torch.cuda.is_available()
    sizes, angles = c4_geometry()
    b = Batch(sizes, sizes, angles)
    images = b.empty_images()
    for i, n in enumerate(sizes):
        b.image(images, i).copy_(phantom(random_ellipses(32, 1000 + i), n))
    b.tune(reps=1)
    sinos = b.radon(images, graph=True)
    bps = b.backprojection(sinos, graph=True)
    torch.cuda.synchronize()
    pick = stratified(sizes, angles)
    print("C4 sample:", [(i, sizes[i], angles[i]) for i in pick], flush=True)
    return torch, b, sizes, angles, images, sinos, bps, pick


def test_c4_sample_covers_every_size(c4):
    torch, b, sizes, angles, images, sinos, bps, pick = c4
    assert len(sizes) == 256 and len(pick) >= 12
    assert sorted({sizes[i] for i in pick}) == SIZES
    assert max(angles[i] for i in pick) >= 1780 and min(angles[i] for i in pick) <= 200


def _angle_blocks(A):
    q = A // 4
    blocks = [(0, 4), (q - 2, q + 2), (2 * q - 2, 2 * q + 2), (A - 4, A)]
    return blocks


def test_c4_radon_vs_oracle(c4):
    torch, b, sizes, angles, images, sinos, bps, pick = c4
    for i in pick:
        n, A = sizes[i], angles[i]
        f = b.image(images, i).cpu().numpy().astype(np.float64)
        got = b.sino(sinos, i).cpu().numpy()
        if n <= 768:
            e = rel_l2(got, O.radon(f, n, A))
            print("radon slice %d (n %d, A %d): rel-L2 %.2e" % (i, n, A, e), flush=True)
            assert e <= REL_L2
        else:
            for j0, j1 in _angle_blocks(A):
                ref = O.radon(f, n, A, j0=j0, j1=j1)[:, j0:j1]
                e = rel_l2(got[:, j0:j1], ref)
                print("radon slice %d (n %d, A %d) angles [%d, %d): rel-L2 %.2e" % (i, n, A, j0, j1, e), flush=True)
                assert e <= REL_L2


def _rows(n):
    return sorted(set([0, 1, 2, n // 2 - 1, n // 2, n - 3, n - 2, n - 1] + list(range(5, n, max(1, n // 12)))))


def test_c4_backprojection_vs_oracle(c4):
    torch, b, sizes, angles, images, sinos, bps, pick = c4
    for i in pick:
        n, A = sizes[i], angles[i]
        s = b.sino(sinos, i).cpu().numpy().astype(np.float64)
        got = b.image(bps, i).cpu().numpy()
        if n <= 512:
            e = rel_l2(got, O.backprojection(s, n))
        else:
            rows = _rows(n)
            e = rel_l2(got[rows], O.backprojection_rows(s, n, rows))
        print("backprojection slice %d (n %d, A %d): rel-L2 %.2e" % (i, n, A, e), flush=True)
        assert e <= REL_L2


def test_c4_osem_8_subsets(c4):
    torch, b, sizes, angles, images, sinos, bps, pick = c4
    from pyraft_amd.device import Plan
    niter, S = 2, 8
    x = torch.ones(b.img_total, dtype=torch.float32, device=sinos.device)
    x, ll = b.em(sinos, niter, S, x=x, return_ll=True, graph=True)
    torch.cuda.synchronize()
    for i in pick:
        n, A = sizes[i], angles[i]
        got = b.image(x, i).cpu().numpy()
        s = b.sino(sinos, i)
        if n <= 768:
            ref, llref = O.osem(s.cpu().numpy().astype(np.float64), n, niter, S)
            e = rel_l2(got, ref)
            print("osem slice %d (n %d, A %d): rel-L2 %.2e vs oracle, LL %s vs %s"
                  % (i, n, A, e, [ll[k][i] for k in range(niter)], llref), flush=True)
            assert e <= REL_L2
            for k in range(niter):
                assert abs(ll[k][i] - llref[k]) <= 1e-6 * abs(llref[k])
        else:
            plain, llp = Plan(n, n, A).em(s.contiguous(), niter, S, return_ll=True)
            e = rel_l2(got, plain.cpu().numpy())
            print("osem slice %d (n %d, A %d): rel-L2 %.2e vs per-slice call" % (i, n, A, e), flush=True)
            assert e <= REL_L2
            for k in range(niter):
                assert abs(ll[k][i] - llp[k]) <= 1e-6 * abs(llp[k])
