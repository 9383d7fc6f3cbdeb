"""Emission EM and OSEM (pyraft/iterative/em.py).

``em`` and ``make_em`` keep pyraft's defaults: the line-traversal (Bresenham) pair,
composed exactly as pyraft/iterative/em.py:109-169 composes it (x <- x * R^T(s / R x)
/ R^T 1 with unguarded numpy divisions), each projection a GPU call.

``make_em(..., projection_method=SLANTSTACK, backprojection_method=SLANTSTACK)``,
``osem`` and ``em_gpu`` run the multiplicative recurrence of raft/raft_em_gpu.cu:41-69,
x <- x * BP_b(s / R_b x) / BP_b(1), with the slant-stack Radon R and the slant-stack
backprojection BP, as fused HIP kernels: the forward projection writes the
guarded ratio (0 where R x <= 0; the reference divides unguarded,
raft/raft_cuda_aux.cu:21) and FP64 log-likelihood partials, the backprojection
applies the update in its epilogue.  OSEM subset b covers the contiguous angle
columns [b*A/S, (b+1)*A/S) (pyraft/iterative/em.py:284-312) on the full
sinogram's angle grid.  Pixels with BP_b(1) = 0 keep their value.  (pyraft's
``osem`` uses the NFFT slice pair, which pyraft_amd does not implement.)
"""
import ctypes
import math

import numpy

from .raftypes import check_legacy, image, libraft
from .sinogram import radon_method

__all__ = ["make_em", "em", "em_gpu", "em_transmission_gpu", "osem"]


def _corners(arr, default_tl, default_br):
    tl = getattr(arr, 'top_left', None) or default_tl
    br = getattr(arr, 'bottom_right', None) or default_br
    return (float(tl[0]), float(tl[1]), float(br[0]), float(br[1]))


def _sino_corners(sino):
    return _corners(sino, (0.0, 1.0), (math.pi, -1.0))


def _plan_for(x_shape, x_corners, sino):
    from .device import Plan
    if x_shape[0] != x_shape[1]:
        raise ValueError("slant-stack EM needs a square image (got %s)" % (tuple(x_shape),))
    return Plan(x_shape[0], sino.shape[0], sino.shape[1], x_corners, _sino_corners(sino))


def _device(a):
    import torch
    return torch.from_numpy(numpy.ascontiguousarray(numpy.asarray(a, dtype=numpy.float32))).cuda()


def _composed_em(sino, shape, projection_method, backprojection_method, weight, **kwargs):
    """pyraft/iterative/em.py:109-142 with pyraft_amd's radon / radon_transpose."""
    from .backprojection import radon_transpose
    from .sinogram import radon
    if weight is None:
        one = image(numpy.ones(sino.shape), top_left=sino.top_left, bottom_right=sino.bottom_right)
        if isinstance(shape, image):
            weight = radon_transpose(one, shape.shape, top_left=shape.top_left, bottom_right=shape.bottom_right,
                                     method=backprojection_method, **kwargs)
        else:
            weight = radon_transpose(one, shape, method=backprojection_method, **kwargs)

    def em_iter(x):
        Ax = radon(x, sino.shape, top_left=sino.top_left, bottom_right=sino.bottom_right, method=projection_method)
        Ax = sino / Ax
        return x * radon_transpose(Ax, x.shape, top_left=x.top_left, bottom_right=x.bottom_right,
                                   method=backprojection_method) / weight

    return em_iter


def make_em(sino, shape, projection_method=radon_method.BRESENHAM, backprojection_method=radon_method.BRESENHAM,
            weight=None, **kwargs):
    """One-iteration EM closure (pyraft/iterative/em.py:109-144).  The slant-stack pair
    runs as one fused device iteration; any other pair is composed as in pyraft."""
    if (projection_method, backprojection_method) != (radon_method.SLANTSTACK, radon_method.SLANTSTACK):
        with numpy.errstate(divide="ignore", invalid="ignore"):
            return _composed_em(sino, shape, projection_method, backprojection_method, weight, **kwargs)
    if weight is not None:
        raise NotImplementedError("make_em: the weight is BP(1) of the slant-stack backprojection")
    if isinstance(shape, image):
        xs, xc = shape.shape, _corners(shape, (-1.0, 1.0), (1.0, -1.0))
    else:
        tmp = image(shape, **kwargs)
        xs, xc = tmp.shape, _corners(tmp, (-1.0, 1.0), (1.0, -1.0))
    plan = _plan_for(xs, xc, sino)
    s_dev = _device(sino)

    def em_iter(x):
        out = plan.em(s_dev, 1, 1, x0=_device(x)).cpu().numpy().astype(numpy.float64)
        return image(out, top_left=(xc[0], xc[1]), bottom_right=(xc[2], xc[3]))

    return em_iter


def em(sino, shape=None, niter=5, **kwargs):
    """EM from ones (or from ``shape`` when it is an image) with pyraft's default
    (line-traversal) pair, pyraft/iterative/em.py:151-169; the default image is
    A x A for an R x A sinogram, as in pyraft."""
    if shape is None:
        shape = [sino.shape[1], sino.shape[1]]
    if isinstance(shape, image):
        if kwargs:
            raise TypeError('Unhandled arguments!')
        x = shape
    else:
        x = image(shape, **kwargs)
        x[:] = 1.0
    step = make_em(sino, x)
    with numpy.errstate(divide="ignore", invalid="ignore"):
        for _ in range(0, int(niter)):
            x = step(x)
    return x


def em_gpu(*args):
    """em_gpu(s, img_size=R, niter=20) -> float32 (pyraft/iterative/em.py:176-206)."""
    if len(args) == 0:
        raise TypeError('Not enough arguments!')
    s = args[0]
    rays, angles = s.shape[0], s.shape[1]
    img_size = int(args[1]) if len(args) > 1 else rays
    niter = int(args[2]) if len(args) > 2 else 20
    sino_buff = numpy.ascontiguousarray(s, dtype=numpy.float32)
    img = numpy.zeros(img_size * img_size, dtype=numpy.float32)
    libraft.raft_em_gpu(img.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                        sino_buff.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), img_size, rays, angles, niter)
    check_legacy("raft_em_gpu")
    return img.reshape([img_size, img_size])


def em_transmission_gpu(*args):
    """em_transmission_gpu(s, flat, dark, img_size=R, niter=20) -> float64 image
    (pyraft/iterative/em.py:212-248): raft_tr_em_gpu's transmission EM from x = 1.  pyraft
    passes the dark field to a function that takes none (raft/raft_transmission_em.cu:16),
    so the call cannot succeed there; here dark is accepted and unused, as that function
    would use it."""
    if len(args) < 3:
        raise TypeError('Not enough arguments!')
    s, f = args[0], args[1]
    rays, angles = s.shape[0], s.shape[1]
    img_size = int(args[3]) if len(args) > 3 else rays
    niter = int(args[4]) if len(args) > 4 else 20
    sino_buff = numpy.ascontiguousarray(s, dtype=numpy.float32)
    flat_buff = numpy.ascontiguousarray(f, dtype=numpy.float32)
    if flat_buff.shape != sino_buff.shape:
        raise ValueError("flat field must have the sinogram's shape %s" % (sino_buff.shape,))
    img = numpy.zeros(img_size * img_size, dtype=numpy.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    libraft.raft_tr_em_gpu(img.ctypes.data_as(fp), sino_buff.ctypes.data_as(fp), flat_buff.ctypes.data_as(fp),
                           img_size, rays, angles, niter)
    check_legacy("raft_tr_em_gpu")
    return img.reshape([img_size, img_size]).astype('float64')


def osem(sino, shape=None, niter=5, nblocks=4, **kwargs):
    """OSEM with ``nblocks`` contiguous angle subsets from x = 1 on an R x R image
    (pyraft/iterative/em.py:265-320; ``shape`` selects another square size)."""
    n = sino.shape[0] if shape is None else int(shape[0])
    x = image((n, n), **kwargs)
    xc = _corners(x, (-1.0, 1.0), (1.0, -1.0))
    plan = _plan_for(x.shape, xc, sino)
    out = plan.em(_device(sino), int(niter), int(nblocks)).cpu().numpy().astype(numpy.float64)
    return image(out, top_left=(xc[0], xc[1]), bottom_right=(xc[2], xc[3]))
