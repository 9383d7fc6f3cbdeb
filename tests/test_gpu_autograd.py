"""Differentiable projector on 64 slices of 512 x 512, 720 angles (FP32).

* dot-product adjoint test through autograd: <R x, y> vs <x, grad_x <R x, y>>,
  FP64 dots, mismatch <= 1e-6;
* central finite differences (h = 1e-3) of f(x) = 0.5 ||R x - y||^2 on 1000 random
  pixels against the autograd gradient (f evaluated in FP64 from the FP32
  projections; R is linear, so R(x +- h e_i) = R x +- h R e_i, and R e_i is the
  projection of a one-hot image);
* a 20-step unrolled least-squares + TV reconstruction with per-step checkpointing:
  its data loss falls, its gradient w.r.t. the TV weight is finite, and the
  checkpointed gradient equals the stored-activation gradient.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import torch
    from pyraft_amd.device import Plan
    assert torch.cuda.is_available()
    torch.manual_seed(5)
    plan = Plan(512, 512, 720)
    return torch, plan


def test_adjoint_dot_product(setup):
    torch, plan = setup
    from pyraft_amd.autograd import radon
    print("adjoint 64x512^2x720", flush=True)
    x = torch.rand((64, 512, 512), device="cuda", requires_grad=True)
    y = torch.rand((64, 512, 720), device="cuda")
    Rx = radon(x, plan)
    lhs = float((Rx.double() * y.double()).sum())
    g, = torch.autograd.grad((Rx * y).sum(), x)
    rhs = float((x.detach().double() * g.double()).sum())
    mism = abs(lhs - rhs) / abs(lhs)
    print("adjoint mismatch", mism)
    assert mism <= 1e-6


def test_finite_differences(setup):
    torch, plan = setup
    from pyraft_amd.autograd import radon
    print("finite differences", flush=True)
    gen = torch.Generator(device="cuda").manual_seed(7)
    x = torch.rand((64, 512, 512), device="cuda", generator=gen, requires_grad=True)
    y = torch.rand((64, 512, 720), device="cuda", generator=gen) * 50.0
    r = radon(x, plan) - y
    loss = 0.5 * (r * r).sum()
    g, = torch.autograd.grad(loss, x)
    r = r.detach()
    rng = np.random.default_rng(11)
    picks = rng.choice(64 * 512 * 512, size=1000, replace=False)
    h = 1e-3
    fd = np.zeros(1000)
    ad = g.reshape(-1)[torch.from_numpy(picks).cuda()].double().cpu().numpy()
    for c0 in range(0, 1000, 250):
        idx = picks[c0:c0 + 250]
        sl, pix = idx // (512 * 512), idx % (512 * 512)
        e = torch.zeros((len(idx), 512, 512), device="cuda")
        e.view(len(idx), -1)[torch.arange(len(idx), device="cuda"), torch.from_numpy(pix).cuda()] = 1.0
        Re = plan.radon(e).double()                       # (B, 512, 720)
        rs = r[torch.from_numpy(sl).cuda()].double()      # residual of each pick's slice
        fp = 0.5 * ((rs + h * Re) ** 2).sum(dim=(1, 2))
        fm = 0.5 * ((rs - h * Re) ** 2).sum(dim=(1, 2))
        fd[c0:c0 + 250] = ((fp - fm) / (2 * h)).cpu().numpy()
    err = np.linalg.norm(fd - ad) / np.linalg.norm(fd)
    print("finite-difference rel error", err)
    assert err <= 1e-5


def test_unrolled_ls_tv_checkpointed(setup):
    torch, plan = setup
    from pyraft_amd.autograd import unrolled_ls_tv
    from pyraft_amd.device import phantom
    from pyraft_amd.phantom import random_ellipses
    print("unrolled LS+TV", flush=True)
    x_true = torch.stack([phantom(random_ellipses(16, s), 512) for s in range(64)])
    y = plan.radon(x_true)
    lam = torch.tensor(1e-3, device="cuda", requires_grad=True)
    x, losses = unrolled_ls_tv(y, plan, steps=20, lam=lam, eps=1e-8, checkpoint=True)
    print("data loss first/last", losses[0], losses[-1])
    assert np.isfinite(losses).all() and losses[-1] < 0.5 * losses[0]
    obj = ((x - x_true) ** 2).mean()
    g_ck, = torch.autograd.grad(obj, lam)
    assert torch.isfinite(g_ck)
    # same gradient without checkpointing (8 slices to keep stored activations small)
    from pyraft_amd.autograd import operator_norm_sq
    alpha = 0.95 / operator_norm_sq(plan)
    assert alpha == 0.95 / operator_norm_sq(plan)   # seeded power iteration
    y8 = y[:8].contiguous()
    xt8 = x_true[:8]
    lam2 = torch.tensor(1e-3, device="cuda", requires_grad=True)
    xa, _ = unrolled_ls_tv(y8, plan, steps=20, step_size=alpha, lam=lam2, checkpoint=True)
    ga, = torch.autograd.grad(((xa - xt8) ** 2).mean(), lam2)
    lam3 = torch.tensor(1e-3, device="cuda", requires_grad=True)
    xb, _ = unrolled_ls_tv(y8, plan, steps=20, step_size=alpha, lam=lam3, checkpoint=False)
    gb, = torch.autograd.grad(((xb - xt8) ** 2).mean(), lam3)
    print("lambda grads", float(ga), float(gb))
    assert abs(float(ga) - float(gb)) <= 1e-3 * abs(float(gb)) + 1e-12
