"""Batched execution of uneven slices and the backprojection kernel variants (GPU).

* every backprojection variant matches the FP64 oracle (rel-L2 <= 1e-5) and the
  autotuner keeps a variant that fits;
* a batch of slices with different n, R, A and corners gives each slice the result
  of its own per-plan call: bit-identical backprojection (same variant, same
  device function; within 1e-6 where the one-slice call divides its angles into
  groups to fill the GPU), Radon within 1e-6 (the batched launch may split sample ranges
  differently) and the oracle within 1e-5;
* batched OSEM matches per-slice OSEM, log-likelihood included;
* the graph-captured form is bit-identical to the direct form on every replay, with
  other GPU work and host synchronisation between the replays.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

REL_L2 = 1e-5
PI = math.pi
DEF_IC = (-1.0, 1.0, 1.0, -1.0)
DEF_SC = (0.0, 1.0, PI, -1.0)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (den if den > 0 else 1.0))


def _geo(ic, sc):
    return dict(img_tl=(ic[0], ic[1]), img_br=(ic[2], ic[3]), sino_tl=(sc[0], sc[1]), sino_br=(sc[2], sc[3]))


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


# n, R, A, img corners, sino corners
SLICES = [
    (64, 64, 45, DEF_IC, DEF_SC),
    (100, 100, 73, DEF_IC, DEF_SC),
    (48, 40, 91, (-0.7, 1.3, 1.2, -0.9), (-PI / 2, 1.1, PI / 2, -0.8)),
    (128, 131, 180, DEF_IC, DEF_SC),
    (64, 64, 45, DEF_IC, DEF_SC),        # same geometry as slice 0 (shared plan)
    (33, 50, 17, DEF_IC, DEF_SC),
    (256, 256, 200, DEF_IC, DEF_SC),
]


def _batch(torch, slices=SLICES):
    from pyraft_amd.batch import Batch
    return Batch([s[0] for s in slices], [s[1] for s in slices], [s[2] for s in slices],
                 img_corners=[s[3] for s in slices], sino_corners=[s[4] for s in slices])


def _images(slices, seed):
    rng = np.random.default_rng(seed)
    return [rng.random((s[0], s[0])).astype(np.float32) for s in slices]


def _sinos(slices, seed):
    rng = np.random.default_rng(seed)
    return [rng.random((s[1], s[2])).astype(np.float32) for s in slices]


@pytest.mark.parametrize("variant", range(6))
def test_bp_variants_match_oracle(torch_mod, variant):
    torch = torch_mod
    from pyraft_amd.device import Plan, bp_variant_shape
    print("bp variant", variant, bp_variant_shape(variant), flush=True)
    for n, R, A, ic, sc in [(100, 100, 73, DEF_IC, DEF_SC), (48, 40, 91, SLICES[2][3], SLICES[2][4]),
                            (256, 256, 180, DEF_IC, DEF_SC)]:
        p = Plan(n, R, A, ic, sc)
        p.set_bp_variant(variant)
        assert p.info()["bp_variant"] == variant
        s = np.random.default_rng(3).random((R, A)).astype(np.float32)
        got = p.backprojection(torch.from_numpy(s).cuda()).cpu().numpy()
        ref = O.backprojection(s.astype(np.float64), n, **_geo(ic, sc))
        err = rel_l2(got, ref)
        print(n, R, A, "rel-L2", err, flush=True)
        assert err <= REL_L2


def test_plan_tune(torch_mod):
    from pyraft_amd.device import Plan
    p = Plan(256, 256, 180)
    v, times = p.tune(reps=2)
    print("tuned variant", v, times, flush=True)
    assert 0 <= v < 6 and v in times and times[v] == min(times.values())
    assert p.info()["bp_variant"] == v
    v2, _ = p.tune()  # cached
    assert v2 == v


def test_batch_radon_matches_per_slice(torch_mod):
    torch = torch_mod
    from pyraft_amd.device import Plan
    b = _batch(torch)
    imgs = _images(SLICES, 21)
    packed = b.pack_images([torch.from_numpy(x).cuda() for x in imgs])
    for angle_major in (False, True):
        print("batch radon, angle_major", angle_major, flush=True)
        out = b.radon(packed, angle_major=angle_major)
        for i, (n, R, A, ic, sc) in enumerate(SLICES):
            got = b.sino(out, i, angle_major).cpu().numpy()
            own = Plan(n, R, A, ic, sc).radon(torch.from_numpy(imgs[i]).cuda(), angle_major=angle_major)
            own = own.cpu().numpy()
            assert rel_l2(got, own) <= 1e-6, i
            ref = O.radon(imgs[i].astype(np.float64), R, A, **_geo(ic, sc))
            assert rel_l2(got.T if angle_major else got, ref) <= REL_L2, i


def test_batch_backprojection_bit_identical(torch_mod):
    torch = torch_mod
    from pyraft_amd.device import Plan
    b = _batch(torch)
    sinos = _sinos(SLICES, 22)
    packed = b.pack_sinos([torch.from_numpy(s).cuda() for s in sinos])
    out = b.backprojection(packed)
    for i, (n, R, A, ic, sc) in enumerate(SLICES):
        print("batch bp slice", i, flush=True)
        got = b.image(out, i).cpu().numpy()
        p = Plan(n, R, A, ic, sc)
        p.set_bp_variant(b.info(i)["bp_variant"])
        own = p.backprojection(torch.from_numpy(sinos[i]).cuda()).cpu().numpy()
        if p.info()["bp_groups"] == 1:
            assert np.array_equal(got, own), i
        else:  # the one-slice launch divides its angles into groups: other FP32 summation order
            assert rel_l2(got, own) <= 1e-6, i
        ref = O.backprojection(sinos[i].astype(np.float64), n, **_geo(ic, sc))
        assert rel_l2(got, ref) <= REL_L2, i


def test_batch_variants_and_tune(torch_mod):
    torch = torch_mod
    b = _batch(torch)
    sinos = _sinos(SLICES, 23)
    packed = b.pack_sinos([torch.from_numpy(s).cuda() for s in sinos])
    base = b.backprojection(packed).clone()
    for v in (1, 2, 5):
        print("batch variant", v, flush=True)
        b.set_bp_variant(v)
        assert all(b.info(i)["bp_variant"] == v for i in range(len(b)))
        got = b.backprojection(packed)
        for i in range(len(b)):
            assert rel_l2(b.image(got, i).cpu().numpy(), b.image(base, i).cpu().numpy()) <= 1e-6, (v, i)
    chosen = b.tune(reps=1)
    print("tuned", chosen, flush=True)
    assert all(0 <= v < 6 for v in chosen)
    # slices of one image size share the tuned variant
    by_n = {}
    for i, v in enumerate(chosen):
        by_n.setdefault(SLICES[i][0], set()).add(v)
    assert all(len(s) == 1 for s in by_n.values())


def test_batch_em_matches_per_slice(torch_mod):
    torch = torch_mod
    from pyraft_amd.device import Plan
    b = _batch(torch)
    niter, S = 3, 4
    imgs = _images(SLICES, 24)
    sinos = []
    for (n, R, A, ic, sc), x in zip(SLICES, imgs):
        sinos.append(Plan(n, R, A, ic, sc).radon(torch.from_numpy(x + 0.1).cuda()))
    packed = b.pack_sinos(sinos)
    print("batch em", flush=True)
    x, ll = b.em(packed, niter, S, return_ll=True)
    for i, (n, R, A, ic, sc) in enumerate(SLICES):
        p = Plan(n, R, A, ic, sc)
        p.set_bp_variant(b.info(i)["bp_variant"])
        own, own_ll = p.em(sinos[i], niter, S, return_ll=True)
        got = b.image(x, i).cpu().numpy()
        err = rel_l2(got, own.cpu().numpy())
        print("slice", i, "rel-L2 vs per-slice", err, flush=True)
        assert err <= 1e-5, i
        for it in range(niter):
            assert abs(ll[it][i] - own_ll[it]) <= 1e-6 * abs(own_ll[it]) + 1e-9, (i, it)


def test_batch_graph_replays_bit_identical(torch_mod):
    torch = torch_mod
    b = _batch(torch)
    imgs = _images(SLICES, 25)
    packed = b.pack_images([torch.from_numpy(x).cuda() for x in imgs])
    eager = b.radon(packed)
    sinos = b.radon(packed).clone()
    eager_bp = b.backprojection(sinos).clone()
    eager_em, eager_ll = b.em(sinos, 2, 3, return_ll=True)
    eager_em = eager_em.clone()
    g = b.empty_sinos()
    gb = b.empty_images()
    x = torch.ones_like(eager_em)
    for rep in range(3):
        # fixed buffers: the first pass captures, the later ones replay
        print("graph replay", rep, flush=True)
        g.zero_()
        b.radon(packed, out=g, graph=True)
        torch.cuda.synchronize()
        assert torch.equal(g, eager)
        gb.zero_()
        b.backprojection(sinos, out=gb, graph=True)
        torch.cuda.synchronize()
        assert torch.equal(gb, eager_bp)
        x.fill_(1.0)
        gx, gll = b.em(sinos, 2, 3, x=x, return_ll=True, graph=True)
        assert torch.equal(gx, eager_em)
        assert gll == eager_ll
        # unrelated eager work between replays
        junk = torch.randn(1 << 20, device="cuda")
        (junk * 2).sum().item()


def test_one_slice_batch_em_is_the_plan_em(torch_mod):
    """A one-slice batch runs the slice's own OSEM kernels: the result and the LL trace
    equal the plan's OSEM bit for bit, eagerly and in graph replays (with eager work
    between the replays)."""
    torch = torch_mod
    from pyraft_amd.device import Plan
    n, R, A = 160, 170, 96
    b = _batch(torch, [(n, R, A, DEF_IC, DEF_SC)])
    p = Plan(n, R, A)
    x0 = torch.from_numpy(_images([(n, R, A, DEF_IC, DEF_SC)], 31)[0] + 0.2).cuda()
    sino = p.radon(x0)
    packed = b.pack_sinos([sino])
    own, own_ll = p.em(sino, 3, 4, return_ll=True)
    print("eager one-slice batch em", flush=True)
    x, ll = b.em(packed, 3, 4, return_ll=True)
    assert torch.equal(b.image(x, 0), own)
    assert [l[0] for l in ll] == own_ll
    xg = torch.ones_like(x)
    for rep in range(3):
        print("graph replay", rep, flush=True)
        xg.fill_(1.0)
        gx, gll = b.em(packed, 3, 4, x=xg, return_ll=True, graph=True)
        assert torch.equal(b.image(gx, 0), own)
        assert [l[0] for l in gll] == own_ll
        junk = torch.randn(1 << 18, device="cuda")
        (junk * 2).sum().item()
