"""Radon transforms (pyraft/sinogram/sino.py).

``radon(img, shape, method=radon_method.SLANTSTACK, **kwargs)`` keeps pyraft's
signature and sinogram layout (rows = rays, columns = angles, default corners
theta in (0, pi), t in (-1, 1); pyraft/sinogram/sino.py:175-212) and calls the
drop-in ``raft_radon_slantstack`` symbol of libraft, which runs on the GPU.

``radon_method.BRESENHAM`` (the default, the line-traversal projector of
raft/raft_radon_bresenham.cpp) and ``radon_method.SLANTSTACK`` run on the GPU; the
Fourier methods raise ``NotImplementedError``.

The arbitrary-mesh forms ``radon_view`` / ``radon_withmesh`` / ``radon_3d_stack`` /
``snapshot`` (pyraft/sinogram/sino.py:220-343) take t and theta from tables.  As in
pyraft, the image passed to them always spans [-1, 1]^2: pyraft's make_RAFT_IMAGE
reaches its ndarray branch before its image branch (pyraft/raftypes.py:427-449), so
an image's own corners never reach the C side.
"""
import ctypes
import math

import numpy

from .raftypes import (RAFT_IMAGE, check_legacy, image, libraft, make_RAFT_IMAGE, make_RAFT_MATRIX,
                       make_RAFT_VECTOR, nthreads)

__all__ = ["radon_method", "radon", "radon_gpu", "radon_view", "radon_withmesh", "radon_3d_stack", "snapshot"]


class radon_method:
    """Methods for projection/backprojection operations (pyraft/sinogram/sino.py:159-167).

    Members are 1-tuples exactly as in pyraft, so identity/equality tests written
    against pyraft keep working."""
    BRESENHAM = 1,
    SLANTSTACK = 2,
    HIERARCHICAL = 3,
    BST = 4,
    ANDERSSON = 5,
    ANDERSSONPB = 6


def _float64_image(img):
    if isinstance(img, image) and img.dtype == numpy.float64:
        return img
    out = image(numpy.asarray(img, dtype=numpy.float64), top_left=tuple(img.top_left),
                bottom_right=tuple(img.bottom_right))
    return out


def radon(img, shape=(367, 180), method=radon_method.BRESENHAM, **kwargs):
    """Radon transform of a pyraft ``image`` (line traversal or slant-stack) or of an
    ellipse description.

    pyraft/sinogram/sino.py:175-212.  With no extent keywords the sinogram spans
    theta in (0, pi) (``theta_extent`` overrides) and t in (-1, 1) (``t_extent``)."""
    if ('top_left' not in kwargs) and ('bottom_right' not in kwargs) and \
       ('x_extent' not in kwargs) and ('y_extent' not in kwargs):
        kwargs['x_extent'] = kwargs.pop('theta_extent', (0.0, math.pi))
        if 't_extent' in kwargs:
            kwargs['y_extent'] = kwargs.pop('t_extent')

    sino = image(shape, **kwargs)
    SINO = make_RAFT_IMAGE(sino, sino.top_left, sino.bottom_right)

    if isinstance(img, image):
        if method == radon_method.SLANTSTACK:
            src = _float64_image(img)
            IMG = make_RAFT_IMAGE(src, src.top_left, src.bottom_right)
            libraft.raft_radon_slantstack(IMG, SINO, ctypes.c_int(nthreads))
            check_legacy("raft_radon_slantstack")
        elif method == radon_method.BRESENHAM:
            src = _float64_image(img)
            IMG = make_RAFT_IMAGE(src, src.top_left, src.bottom_right)
            libraft.raft_radon_bresenham(IMG, SINO, ctypes.c_int(nthreads))
            check_legacy("raft_radon_bresenham")
        else:
            raise TypeError('Unsupported method for Radon Transform!')
        return sino

    if isinstance(img, numpy.ndarray):
        desc = numpy.ascontiguousarray(img, dtype=numpy.float64)
        libraft.raft_radon_fromdesc(SINO, make_RAFT_MATRIX(desc))
        check_legacy("raft_radon_fromdesc")
        return sino

    raise TypeError('radon needs a pyraft image or an ellipse description')


def radon_gpu(*args):
    """radon_gpu(f, rays=f.shape[0], angles=360) (pyraft/sinogram/sino.py:350-387).

    Host float32 round trip through ``raft_radon_slantstack_gpu`` with the default
    geometry; returns a float64 ``image`` with corners (0, 1), (pi, -1)."""
    if len(args) == 0:
        raise TypeError('Not enough arguments!')
    f = args[0]
    rays = f.shape[0] if len(args) < 2 else int(args[1])
    angles = 360 if len(args) < 3 else int(args[2])
    img_size = f.shape[0]
    f_buff = numpy.ascontiguousarray(f, dtype=numpy.float32)
    sino = numpy.zeros(rays * angles, dtype=numpy.float32)
    libraft.raft_radon_slantstack_gpu(sino.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                      f_buff.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                      img_size, rays, angles)
    check_legacy("raft_radon_slantstack_gpu")
    out = image(sino.reshape((rays, angles)).astype(numpy.float64))
    out.top_left = (0, 1.0)
    out.bottom_right = (numpy.pi, -1.0)
    return out


def _vector(v):
    return numpy.ascontiguousarray(v, dtype=numpy.float64).reshape(-1)


def _default_t(f):
    # pyraft/sinogram/sino.py:229-233: t spans the image's x extent, one ray per column
    tl = getattr(f, 'top_left', (-1.0, 1.0))
    br = getattr(f, 'bottom_right', (1.0, -1.0))
    return numpy.linspace(tl[0], br[0], f.shape[1])


def radon_view(*args):
    """radon_view(f, theta, t=linspace(f.top_left[0], f.bottom_right[0], f.shape[1]))
    (pyraft/sinogram/sino.py:220-247): the ray sums of one angle, a float64 vector of len(t)."""
    if len(args) == 0:
        raise TypeError('Not enough arguments!')
    f = args[0]
    theta = args[1]
    t = _vector(args[2]) if len(args) > 2 else _default_t(f)
    view = numpy.zeros(t.shape)
    F = make_RAFT_IMAGE(numpy.asarray(f, dtype=numpy.float64))
    libraft.raft_radon_slantstack_view(make_RAFT_VECTOR(view), F, make_RAFT_VECTOR(t), ctypes.c_double(theta))
    check_legacy("raft_radon_slantstack_view")
    return view


def radon_withmesh(*args):
    """radon_withmesh(f, theta, t=linspace(f.top_left[0], f.bottom_right[0], f.shape[1]))
    (pyraft/sinogram/sino.py:253-281): image of shape (len(t), len(theta)), element
    (k, j) = ray sum at t[k], theta[j]."""
    if len(args) == 0:
        raise TypeError('Not enough arguments!')
    f = args[0]
    theta = _vector(args[1])
    t = _vector(args[2]) if len(args) > 2 else _default_t(f)
    sino = image((len(t), len(theta)))
    SINO = make_RAFT_IMAGE(sino, sino.top_left, sino.bottom_right)
    F = make_RAFT_IMAGE(numpy.asarray(f, dtype=numpy.float64))
    libraft.raft_radon_slantstack_withmesh(F, SINO, make_RAFT_VECTOR(t), make_RAFT_VECTOR(theta),
                                           ctypes.c_int(nthreads))
    check_legacy("raft_radon_slantstack_withmesh")
    return sino


def _stack_radon(f, t, theta):
    """Slices f[:, i, :] for i < f.shape[0], all in one launch of one mesh plan.
    Returns (f.shape[0], len(theta), len(t)) float64 (angle-major per slice)."""
    import torch

    from .device import Plan
    f = numpy.asarray(f)
    N = f.shape[0]
    if f.ndim != 3 or f.shape[1] < N or f.shape[0] != f.shape[2]:
        raise ValueError('expected a volume f with f.shape[0] == f.shape[2] <= f.shape[1] (got %s)'
                         % (f.shape,))
    slices = numpy.ascontiguousarray(f.transpose(1, 0, 2)[:N], dtype=numpy.float32)
    plan = Plan.mesh(f.shape[0], t, theta)
    out = plan.radon(torch.from_numpy(slices).to(plan.device), angle_major=True)
    return out.cpu().numpy().astype(numpy.float64)


def radon_3d_stack(*args):
    """radon_3d_stack(f, theta=linspace(0, pi, 180), t=linspace(-1, 1, f.shape[1]))
    (pyraft/sinogram/sino.py:287-311): S[:, i, :] = radon_withmesh(f[:, i, :], theta, t).T
    for i < f.shape[0]; S has shape (len(theta), f.shape[0], len(t)).  The slices run as
    one batched device launch."""
    if len(args) == 0:
        raise TypeError('Not enough arguments!')
    f = args[0]
    theta = _vector(args[1]) if len(args) > 1 else numpy.linspace(0, numpy.pi, 180)
    t = _vector(args[2]) if len(args) > 2 else numpy.linspace(-1, 1, f.shape[1])
    r = _stack_radon(f, t, theta)          # (N, V, R)
    return numpy.ascontiguousarray(r.transpose(1, 0, 2))


def snapshot(*args):
    """snapshot(f, theta=0, t=linspace(-1, 1, f.shape[1])) (pyraft/sinogram/sino.py:317-343):
    snap[k, :] = radon_view(f[:, k, :], theta, t) for k < f.shape[0], an (N, N) array
    (len(t) must be N).  As in pyraft, theta is read only from a two-argument call:
    snapshot(f, theta, t) projects at theta = 0 (pyraft/sinogram/sino.py:325-330)."""
    if len(args) == 0:
        raise TypeError('Not enough arguments!')
    f = numpy.asarray(args[0])
    theta = 0.0
    t = numpy.linspace(-1.0, 1.0, f.shape[1])
    if len(args) == 2:
        theta = float(args[1])
    if len(args) > 2:
        t = _vector(args[2])
    N = f.shape[0]
    if len(t) != N:
        raise ValueError('snapshot: len(t) = %d does not match the (%d, %d) output' % (len(t), N, N))
    r = _stack_radon(f, t, [theta])        # (N, 1, R)
    return numpy.ascontiguousarray(r[:, 0, :])
