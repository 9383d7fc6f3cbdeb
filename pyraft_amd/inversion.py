"""Filtered backprojection (pyraft/inversion/inversion.py).

``fbp(sino)`` is ramp filter + slant-stack backprojection onto an N x N image with
N = sino.shape[0] (the reference's CPU ``fbp`` asks for an undefined method,
pyraft/inversion/inversion.py:22; this is the composition its ``fbp_gpu`` performs).
``fbp_gpu(s[, img_size])`` keeps the reference's float32 return type
(pyraft/inversion/inversion.py:32-61).
"""
import math

import numpy

from .raftypes import image

__all__ = ["fbp", "fbp_gpu"]


def _to_device(sino):
    import torch
    return torch.from_numpy(numpy.ascontiguousarray(numpy.asarray(sino, dtype=numpy.float32))).cuda()


def _corners(sino):
    tl = getattr(sino, 'top_left', None) or (0.0, 1.0)
    br = getattr(sino, 'bottom_right', None) or (math.pi, -1.0)
    return (float(tl[0]), float(tl[1]), float(br[0]), float(br[1]))


def fbp(*args, **kwargs):
    """fbp(sino[, n]) -> float64 ``image`` (n x n, default n = sino.shape[0])."""
    if len(args) == 0:
        raise TypeError("fbp needs a sinogram")
    from .device import Plan
    sino = args[0]
    n = int(args[1]) if len(args) > 1 else int(sino.shape[0])
    img = image((n, n), **kwargs)
    ic = (img.top_left[0], img.top_left[1], img.bottom_right[0], img.bottom_right[1])
    plan = Plan(n, sino.shape[0], sino.shape[1], ic, _corners(sino))
    img[...] = plan.fbp(_to_device(sino)).cpu().numpy()
    return img


def fbp_gpu(*args):
    """Ramp filter with the sinogram's own t grid, then backprojection with the
    default geometry (image [-1,1]^2, theta in [0, pi), t in [-1, 1]); float32 result."""
    if len(args) == 0:
        raise TypeError('Not enough arguments!')
    from .device import Plan, filter_rows, ramp_kernel
    import torch
    s = args[0]
    rays, angles = int(s.shape[0]), int(s.shape[1])
    img_size = int(args[1]) if len(args) > 1 else rays
    sc = _corners(s)
    plan = Plan(img_size, rays, angles)
    rows = _to_device(numpy.asarray(s, dtype=numpy.float32).T)
    g = torch.from_numpy(ramp_kernel(rays, sc[3], sc[1])).cuda()
    filt = filter_rows(rows, g)
    return plan.backprojection(filt, angle_major=True).cpu().numpy()
