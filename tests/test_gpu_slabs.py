"""Radon launches of few blocks (a few angles: OSEM subsets) divide the sample rows of
every (angle block, ray group) into slabs taken by separate blocks, then add the slabs'
partial sums in a finishing pass.  These cases run that form over several slices at once,
with a ragged last slab, in the sinogram, angle-major and angle-range forms and in
OSEM, each slice against the FP64 oracle (rel-L2 <= 1e-5) and the OSEM log-likelihood
against the oracle's sum over the slices.
"""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5

# n, R, A, slices: launches of few blocks of (8 angles x 64 rays), with slabs pinned at 32
# rows (96 rows: 3 slabs; 200 rows: 7 slabs, the last of 8 rows)
CASES = [(96, 100, 9, 3), (200, 150, 16, 2), (130, 70, 5, 4)]


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


@pytest.mark.parametrize("case", CASES, ids=lambda c: "n%d_R%d_A%d_s%d" % c)
def test_radon_row_slabs_several_slices(torch_mod, case):
    torch = torch_mod
    from pyraft_amd.device import Plan
    n, R, A, ns = case
    print("row slabs", case, flush=True)
    rng = np.random.default_rng(n + R + A)
    x = rng.random((ns, n, n)).astype(np.float32)
    p = Plan(n, R, A)
    p.set_radon_slab_rows(32)  # the divided form whatever the tuner would pick
    xd = torch.from_numpy(x).cuda()
    got = p.radon(xd).cpu().numpy()
    got_t = p.radon(xd, angle_major=True).cpu().numpy()
    j0, j1 = 1, A - 1
    got_r = p.radon(xd, angles=(j0, j1)).cpu().numpy()
    for i in range(ns):
        ref = O.radon(x[i].astype(np.float64), R, A)
        e = rel_l2(got[i], ref)
        print("slice", i, "rel-L2", e, flush=True)
        assert e <= REL_L2, i
        assert np.array_equal(got_t[i].T, got[i]), i
        assert rel_l2(got_r[i][:, j0:j1], ref[:, j0:j1]) <= REL_L2, i


@pytest.mark.parametrize("case", CASES[:2], ids=lambda c: "n%d_R%d_A%d_s%d" % c)
def test_osem_row_slabs_several_slices(torch_mod, case):
    torch = torch_mod
    from pyraft_amd.device import Plan
    n, R, A, ns = case
    rng = np.random.default_rng(7 * n + A)
    x = rng.random((ns, n, n))
    s = np.stack([O.radon(x[i], R, A) for i in range(ns)]) + 0.5
    p = Plan(n, R, A)
    p.set_radon_slab_rows(32)
    got, ll = p.em(torch.from_numpy(s.astype(np.float32)).cuda(), 3, nsubsets=3, return_ll=True)
    got = got.cpu().numpy()
    ll_ref = np.zeros(3)
    for i in range(ns):
        ref, lli = O.osem(s[i].astype(np.float32).astype(np.float64), n, 3, nsubsets=3)
        e = rel_l2(got[i], ref)
        print("slice", i, "rel-L2", e, flush=True)
        assert e <= REL_L2, i
        ll_ref += np.asarray(lli)
    for it in range(3):
        assert abs(ll[it] - ll_ref[it]) <= 1e-6 * abs(ll_ref[it]), (it, ll[it], ll_ref[it])
