"""Full-size workloads on the GPU.

2048 x 2048, 1500 angles, 2048 bins (the benchmark workload): Radon columns and
backprojections over contiguous angle blocks (including theta = 0, the knife-edge
block around pi/2 and the last angles) against the FP64 oracle, the full-size
adjoint identity, and the FBP row filter.  1024 x 1024, 720 angles, 16 subsets:
two OSEM iterations against the oracle's recurrence.
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
def big():
    import torch
    from pyraft_amd.device import Plan, phantom
    from pyraft_amd.phantom import random_ellipses
    assert torch.cuda.is_available()
    plan = Plan(2048, 2048, 1500)
    f = phantom(random_ellipses(32, 2), 2048)
    s = plan.radon(f)
    return torch, plan, f, s


BLOCKS = [(0, 6), (372, 378), (746, 754), (1120, 1126), (1494, 1500)]


def test_radon_2048_blocks(big):
    torch, plan, f, s = big
    fx = f.cpu().numpy().astype(np.float64)
    sg = s.cpu().numpy()
    print("radon 2048 blocks", flush=True)
    for j0, j1 in BLOCKS:
        ref = O.radon(fx, 2048, 1500, j0=j0, j1=j1)[:, j0:j1]
        e = rel_l2(sg[:, j0:j1], ref)
        print(j0, j1, e)
        assert e <= REL_L2


def test_backprojection_2048_blocks(big):
    torch, plan, f, s = big
    sg = s.cpu().numpy().astype(np.float64)
    print("backprojection 2048 blocks", flush=True)
    for j0, j1 in BLOCKS[1:4]:
        got = plan.backprojection(s, angles=(j0, j1)).cpu().numpy()
        ref = O.backprojection(sg, 2048, j0=j0, j1=j1)
        e = rel_l2(got, ref)
        print(j0, j1, e)
        assert e <= REL_L2


def test_adjoint_2048(big):
    torch, plan, f, s = big
    y = torch.rand((2048, 1500), device="cuda")
    x = torch.rand((2048, 2048), device="cuda")
    lhs = float((plan.radon(x).double() * y.double()).sum())
    rhs = float((x.double() * plan.radon_transpose(y).double()).sum())
    print("adjoint 2048", abs(lhs - rhs) / abs(lhs))
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


def test_fbp_filter_2048(big):
    torch, plan, f, s = big
    got = plan.fbp_filter(s).cpu().numpy()
    ref = O.lowpassino_fbp(s.cpu().numpy().astype(np.float64))
    e = rel_l2(got, ref)
    print("filter 2048", e)
    assert e <= REL_L2


def test_osem_1024_16_subsets():
    import torch
    from pyraft_amd.device import Plan, phantom
    from pyraft_amd.phantom import random_ellipses
    n, A, S = 1024, 720, 16
    plan = Plan(n, n, A)
    p = plan.radon(phantom(random_ellipses(32, 3), n))
    p = p * (1e3 / p.mean())
    s = torch.from_numpy(np.random.default_rng(30).poisson(p.cpu().numpy().astype(np.float64))
                         .astype(np.float32)).cuda()
    x, ll = plan.em(s, 2, S, return_ll=True)
    ref, llref = O.osem(s.cpu().numpy().astype(np.float64), n, 2, S)
    e = rel_l2(x.cpu().numpy(), ref)
    print("osem 1024/720/16 rel-L2", e, ll, llref)
    assert e <= REL_L2
    for a, b in zip(ll, llref):
        assert abs(a - b) <= 1e-6 * abs(b)
