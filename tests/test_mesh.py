"""Slant-stack Radon on arbitrary (t, theta) tables: raft_radon_slantstack_withmesh,
raft_radon_slantstack_view and pyraft's radon_withmesh / radon_view / radon_3d_stack /
snapshot (pyraft/sinogram/sino.py:220-343).

CPU (oracle and host plan logic):
* the oracle reproduces the reference's own C bit for bit on the golden mesh
  vectors and, where oracle/_ref is built, on fresh random tables;
* uniform tables t_k = tmin + k*dt, theta_j = th0 + j*dth reduce the mesh form to
  the uniform-grid oracle bit for bit (the mesh code shares no arithmetic path with
  the uniform one, so a transposed index or a wrong table read fails here);
* the library's host mesh plan gives every ray the exact sample range of the
  reference predicate and a Q32.32 start coordinate within 1e-6 grid units.

GPU (through the C ABI): the golden vectors and larger random meshes within
rel-L2 1e-5 of the FP64 oracle, uniform tables on a mesh plan against the uniform
plan at 2048^2 x 1500, and the 3-D stack / snapshot helpers against the oracle.
"""
import ctypes
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "slantstack_golden.npz")
MESH_CASES = ("mesh_n33_sorted", "mesh_n40_unsorted", "mesh_n24_fov")
DEF_IC = (-1.0, 1.0, 1.0, -1.0)
SPECIAL_THETA = [0.0, math.pi / 4, math.pi / 2, 3 * math.pi / 4, math.pi, -math.pi / 2, 2 * math.pi]
REL_L2 = 1e-5


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


def _tables(seed, R, A, sort=True, spread=1.6):
    rng = np.random.default_rng(seed)
    t = rng.uniform(-spread, spread, R)
    if sort:
        t.sort()
    theta = np.concatenate([rng.uniform(-4.0, 7.0, max(0, A - len(SPECIAL_THETA))), SPECIAL_THETA])[:A]
    return t, theta


def _uniform_tables(R, A, sc=(0.0, 1.0, math.pi, -1.0)):
    # the reference's grid: theta_j = th0 + j*(thf - th0)/A, t_k = tmin + k*(tmax - tmin)/(R - 1)
    # (raft/raft_radon_slantstack.c:154-172)
    dth = (sc[2] - sc[0]) / A
    dt = (sc[1] - sc[3]) / (R - 1)
    return np.array([sc[3] + k * dt for k in range(R)]), np.array([sc[0] + j * dth for j in range(A)])


# ---------------------------------------------------------------- CPU


def test_golden_mesh_vectors(golden):
    for c in MESH_CASES:
        ic = tuple(golden[c + "/corners"])
        img, t, theta = golden[c + "/img"], golden[c + "/t"], golden[c + "/theta"]
        got = O.radon_withmesh(img, t, theta, img_tl=ic[:2], img_br=ic[2:], nthreads=3)
        assert np.array_equal(got, golden[c + "/radon"]), c
        view = O.radon_view(img, t, float(theta[0]), img_tl=ic[:2], img_br=ic[2:])
        assert np.array_equal(view, golden[c + "/view"]), c
        # a single view is one column of the mesh transform
        assert np.array_equal(view, got[:, 0]), c


@pytest.mark.skipif(not O.ref_available(), reason="reference checker oracle/_ref not built here")
@pytest.mark.parametrize("n,R,A,sort", [(31, 40, 37, True), (64, 61, 30, False), (20, 7, 11, True)])
def test_mesh_against_reference_c(n, R, A, sort):
    t, theta = _tables(n * 1000 + R, R, A, sort)
    img = np.random.default_rng(n).random((n, n))
    ref = O.ref_radon_withmesh(img, t, theta, nthreads=1)
    assert np.array_equal(O.radon_withmesh(img, t, theta, nthreads=4), ref)
    assert np.array_equal(O.radon_view(img, t, float(theta[3])), O.ref_radon_view(img, t, float(theta[3])))


@pytest.mark.parametrize("n,R,A,sc", [(64, 64, 45, (0.0, 1.0, math.pi, -1.0)), (33, 40, 7, (0.3, 1.1, 2.5, -0.8)),
                                      (48, 31, 90, (-math.pi / 2, 1.0, math.pi / 2, -1.0))])
def test_uniform_tables_reduce_to_uniform_grid(n, R, A, sc):
    img = np.random.default_rng(7).random((n, n))
    t, theta = _uniform_tables(R, A, sc)
    uni = O.radon(img, R, A, sino_tl=sc[:2], sino_br=sc[2:])
    assert np.array_equal(O.radon_withmesh(img, t, theta), uni)


@pytest.fixture(scope="module")
def lib():
    from pyraft_amd.raftypes import libraft
    return libraft


def host_mesh_plan(lib, n, t, theta, ic=DEF_IC):
    R, A = len(t), len(theta)
    b = np.zeros((A, R, 2), np.int32)
    alpha = np.zeros((A, R), np.int64)
    dp = ctypes.POINTER(ctypes.c_double)
    t = np.ascontiguousarray(t, dtype=np.float64)
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    rc = lib.raft_ss_plan_host_bounds_mesh(n, (ctypes.c_double * 4)(*ic), R, t.ctypes.data_as(dp), A,
                                           theta.ctypes.data_as(dp), b.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                           alpha.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
    return rc, b, alpha


@pytest.mark.parametrize("n,R,A,sort,ic", [(33, 29, 23, True, DEF_IC), (64, 80, 40, False, DEF_IC),
                                           (50, 45, 30, True, (-0.7, 1.3, 1.2, -0.9)),
                                           (129, 100, 60, False, (1.0, 1.0, -1.0, -1.0))])
def test_host_mesh_plan_matches_reference_predicate(lib, n, R, A, sort, ic):
    t, theta = _tables(n + R + A, R, A, sort)
    rc, b, alpha = host_mesh_plan(lib, n, t, theta, ic)
    assert rc == 0
    mask, coord = O.radon_sample_map_mesh(n, t, theta, ic[:2], ic[2:])
    s = np.arange(n)
    irregular = b[..., 1] == -2
    inrange = (s >= b[..., 0][..., None]) & (s <= b[..., 1][..., None])
    assert not np.any((inrange != mask) & ~irregular[..., None])
    for j, k in np.argwhere(irregular):
        idx = np.nonzero(mask[j, k])[0]
        assert idx.size and idx[-1] - idx[0] + 1 != idx.size
    # Q32.32 start coordinate of every ray within 1e-6 grid units of the reference's
    assert np.all(np.abs(alpha / 2.0 ** 32 - coord[..., 0]) < 1e-6)
    # the tables hold rays that miss the image and rays that cross it
    assert np.any(mask.any(axis=2)) and not np.all(mask.any(axis=2))


def test_host_mesh_plan_rejects_bad_tables(lib):
    t, theta = _tables(1, 8, 8)
    rc, _, _ = host_mesh_plan(lib, 16, t[:1], theta)
    assert rc != 0 and b"nrays" in lib.raft_ss_last_error()
    bad = t.copy()
    bad[2] = np.nan
    rc, _, _ = host_mesh_plan(lib, 16, bad, theta)
    assert rc != 0 and b"non-finite" in lib.raft_ss_last_error()
    rc, _, _ = host_mesh_plan(lib, 16, t * 1e12, theta)
    assert rc != 0


# ---------------------------------------------------------------- GPU


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.mark.gpu
def test_gpu_dropins_match_golden(torch_mod, golden):
    import pyraft_amd as P
    from pyraft_amd.raftypes import make_RAFT_IMAGE, make_RAFT_VECTOR, libraft, check_legacy
    for c in MESH_CASES:
        print("mesh golden", c, flush=True)
        ic = tuple(golden[c + "/corners"])
        img, t, theta = golden[c + "/img"], golden[c + "/t"], golden[c + "/theta"]
        ref = golden[c + "/radon"]
        sino = np.zeros(ref.shape)
        libraft.raft_radon_slantstack_withmesh(make_RAFT_IMAGE(img, ic[:2], ic[2:]), make_RAFT_IMAGE(sino),
                                               make_RAFT_VECTOR(t), make_RAFT_VECTOR(theta), 4)
        check_legacy("raft_radon_slantstack_withmesh")
        assert rel_l2(sino, ref) <= REL_L2, c
        view = np.zeros(len(t))
        libraft.raft_radon_slantstack_view(make_RAFT_VECTOR(view), make_RAFT_IMAGE(img, ic[:2], ic[2:]),
                                           make_RAFT_VECTOR(t), ctypes.c_double(float(theta[0])))
        check_legacy("raft_radon_slantstack_view")
        assert rel_l2(view, golden[c + "/view"]) <= REL_L2, c
        if ic == DEF_IC:
            # pyraft's wrappers (default image corners)
            s2 = P.radon_withmesh(P.image(img), theta, t)
            assert s2.shape == ref.shape and rel_l2(s2, ref) <= REL_L2, c
            assert rel_l2(P.radon_view(P.image(img), float(theta[0]), t), golden[c + "/view"]) <= REL_L2, c


@pytest.mark.gpu
@pytest.mark.parametrize("n,R,A,sort", [(256, 300, 200, True), (200, 257, 129, False), (97, 64, 1, True)])
def test_gpu_mesh_plan_matches_oracle(torch_mod, n, R, A, sort):
    torch = torch_mod
    from pyraft_amd.device import Plan
    print("mesh plan", n, R, A, sort, flush=True)
    t, theta = _tables(n * 7 + A, R, A, sort)
    rng = np.random.default_rng(n)
    imgs = rng.random((2, n, n)).astype(np.float32)
    p = Plan.mesh(n, t, theta)
    assert p.is_mesh and p.info()["nrays"] == R and p.info()["nangles"] == A
    x = torch.from_numpy(imgs).cuda()
    got = p.radon(x).cpu().numpy()
    got_t = p.radon(x, angle_major=True).cpu().numpy()
    for i in range(2):
        ref = O.radon_withmesh(imgs[i].astype(np.float64), t, theta)
        assert rel_l2(got[i], ref) <= REL_L2, i
        assert rel_l2(got_t[i].T, ref) <= REL_L2, i
    # operators that need the uniform grid refuse a mesh plan
    with pytest.raises(Exception):
        p.backprojection(torch.zeros((R, A), device="cuda"))


@pytest.mark.gpu
def test_gpu_mesh_uniform_tables_full_size(torch_mod):
    """Uniform tables on a mesh plan against the uniform plan at 2048^2 x 1500: the
    two paths build their trajectories differently (per-ray table vs alpha0 + k*gamma)."""
    torch = torch_mod
    from pyraft_amd.device import Plan
    n, R, A = 2048, 2048, 1500
    t, theta = _uniform_tables(R, A)
    x = torch.rand((n, n), device="cuda", generator=torch.Generator("cuda").manual_seed(5))
    uni = Plan(n, R, A).radon(x)
    mesh = Plan.mesh(n, t, theta).radon(x)
    err = float(torch.linalg.norm((uni - mesh).double()) / torch.linalg.norm(uni.double()))
    print("mesh vs uniform rel-L2", err, flush=True)
    assert err <= 1e-6


@pytest.mark.gpu
def test_gpu_radon_3d_stack_and_snapshot(torch_mod):
    import pyraft_amd as P
    rng = np.random.default_rng(11)
    N = 24
    f = rng.random((N, N + 3, N))
    theta = np.concatenate([np.linspace(0, np.pi, 17), [np.pi / 2]])
    t = np.linspace(-1.2, 1.2, 31)
    S = P.radon_3d_stack(f, theta, t)
    assert S.shape == (len(theta), N, len(t))
    ref = np.stack([O.radon_withmesh(f[:, i, :], t, theta).T for i in range(N)], axis=1)
    assert rel_l2(S, ref) <= REL_L2
    S0 = P.radon_3d_stack(f[:, :N, :])           # default theta (180 in [0, pi]) and t
    ref0 = np.stack([O.radon_withmesh(f[:, i, :], np.linspace(-1, 1, N), np.linspace(0, np.pi, 180)).T
                     for i in range(N)], axis=1)
    assert rel_l2(S0, ref0) <= REL_L2
    # f.shape[1] != N: the default t has f.shape[1] rays, pyraft's (N, N) output needs N
    with pytest.raises(ValueError):
        P.snapshot(f, 0.7)
    g = f[:, :N, :]
    snap = P.snapshot(g, 0.7)
    ref_snap = np.stack([O.radon_view(g[:, k, :], np.linspace(-1.0, 1.0, N), 0.7) for k in range(N)])
    assert snap.shape == (N, N) and rel_l2(snap, ref_snap) <= REL_L2
    # three arguments: pyraft keeps theta = 0
    tt = np.linspace(-0.9, 0.9, N)
    snap3 = P.snapshot(g, 0.7, tt)
    ref3 = np.stack([O.radon_view(g[:, k, :], tt, 0.0) for k in range(N)])
    assert rel_l2(snap3, ref3) <= REL_L2
