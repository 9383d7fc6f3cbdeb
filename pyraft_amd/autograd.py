"""Differentiable slant-stack projector.

``radon(x, plan)`` is a ``torch.autograd.Function`` whose forward is the HIP
slant-stack Radon and whose backward is the HIP exact matched transpose
(``raft_ss_radon_transpose``), so gradients of any loss of R x are exact to FP32
rounding.  ``radon_transpose`` is the dual Function (forward R^T, backward R), which
makes R^T usable inside differentiable graphs (e.g. unrolled reconstructions whose
update contains the data-term gradient R^T (R x - y)).

``unrolled_ls_tv`` runs T steps of gradient descent on
    0.5 ||R x - y||^2 + lam * TV_eps(x),   TV_eps(x) = sum sqrt(|grad x|^2 + eps)
with every step wrapped in ``torch.utils.checkpoint`` so that backpropagating
through the unrolled reconstruction stores one image per step and recomputes the
step's internals (two projector launches) during the backward pass.
"""
import torch
import torch.utils.checkpoint

from .device import Plan

__all__ = ["radon", "radon_transpose", "RadonOperator", "tv_smooth", "operator_norm_sq", "unrolled_ls_tv"]


class _Radon(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan, angle_major):
        ctx.plan = plan
        ctx.angle_major = angle_major
        return plan.radon(x.contiguous(), angle_major=angle_major)

    @staticmethod
    def backward(ctx, grad):
        gx = ctx.plan.radon_transpose(grad.contiguous(), angle_major=ctx.angle_major)
        return gx, None, None


class _RadonT(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, plan, angle_major):
        ctx.plan = plan
        ctx.angle_major = angle_major
        return plan.radon_transpose(y.contiguous(), angle_major=angle_major)

    @staticmethod
    def backward(ctx, grad):
        gy = ctx.plan.radon(grad.contiguous(), angle_major=ctx.angle_major)
        return gy, None, None


def radon(x, plan, angle_major=False):
    """R x for x of shape (..., N, N); differentiable (backward = exact transpose)."""
    return _Radon.apply(x, plan, angle_major)


def radon_transpose(y, plan, angle_major=False):
    """R^T y for y of shape (..., R, A) (or (..., A, R)); differentiable (backward = R)."""
    return _RadonT.apply(y, plan, angle_major)


class RadonOperator(torch.nn.Module):
    """nn.Module wrapper around a geometry plan."""

    def __init__(self, n, nrays, nangles, angle_major=False, **plan_kwargs):
        super().__init__()
        self.plan = Plan(n, nrays, nangles, **plan_kwargs)
        self.angle_major = angle_major

    def forward(self, x):
        return radon(x, self.plan, self.angle_major)

    def transpose(self, y):
        return radon_transpose(y, self.plan, self.angle_major)


def tv_smooth(x, eps=1e-8):
    """Isotropic smoothed total variation of (..., N, N) images (forward differences,
    replicated boundary), summed over all pixels and slices."""
    dx = torch.diff(x, dim=-1, append=x[..., :, -1:])
    dy = torch.diff(x, dim=-2, append=x[..., -1:, :])
    return torch.sqrt(dx * dx + dy * dy + eps).sum()


def _tv_grad(x, eps):
    with torch.enable_grad():
        xx = x.detach().requires_grad_(True) if not x.requires_grad else x
        g, = torch.autograd.grad(tv_smooth(xx, eps), xx, create_graph=x.requires_grad)
    return g


def operator_norm_sq(plan, iters=12, angle_major=False, device=None, seed=0):
    """||R||_2^2 by power iteration on R^T R (a lower bound that converges from below),
    from a seeded random start so that repeated calls return the same value."""
    dev = torch.device(device) if device is not None else plan.device
    with torch.no_grad():
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(seed))
        v = torch.rand((plan.n, plan.n), device=dev, dtype=torch.float32, generator=gen)
        lam = 0.0
        for _ in range(iters):
            v = v / v.norm()
            w = plan.radon_transpose(plan.radon(v, angle_major=angle_major), angle_major=angle_major)
            lam = float((w * v).sum())
            v = w
    return lam


def unrolled_ls_tv(y, plan, steps=20, step_size=None, lam=1e-3, eps=1e-8, x0=None, checkpoint=True,
                   angle_major=False):
    """Unrolled gradient descent on 0.5||Rx - y||^2 + lam*TV_eps(x).

    ``step_size`` and ``lam`` may be tensors requiring grad (learned unrolling); the
    returned reconstruction is differentiable with respect to them, to ``y`` and to
    ``x0``.  Returns (x_T, list of per-step data losses as floats)."""
    lead = y.shape[:-2]
    if x0 is None:
        x0 = torch.zeros(lead + (plan.n, plan.n), dtype=y.dtype, device=y.device)
    if step_size is None:
        step_size = 0.95 / operator_norm_sq(plan, angle_major=angle_major, device=y.device)
    losses = []

    def step(x, y_, alpha, lam_):
        r = radon(x, plan, angle_major) - y_
        g = radon_transpose(r, plan, angle_major)
        return x - alpha * (g + lam_ * _tv_grad(x, eps))

    x = x0
    alpha = step_size if torch.is_tensor(step_size) else torch.tensor(float(step_size), device=y.device)
    lam_t = lam if torch.is_tensor(lam) else torch.tensor(float(lam), device=y.device)
    for _ in range(steps):
        if checkpoint and torch.is_grad_enabled():
            x = torch.utils.checkpoint.checkpoint(step, x, y, alpha, lam_t, use_reentrant=False)
        else:
            x = step(x, y, alpha, lam_t)
        with torch.no_grad():
            losses.append(0.5 * float((plan.radon(x.detach().contiguous(), angle_major=angle_major) - y).double()
                                      .pow(2).sum()))
    return x, losses
