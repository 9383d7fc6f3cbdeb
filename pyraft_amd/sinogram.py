"""Radon transforms (pyraft/sinogram/sino.py).

``radon(img, shape, method=radon_method.SLANTSTACK, **kwargs)`` keeps pyraft's
signature and sinogram layout (rows = rays, columns = angles, default corners
theta in (0, pi), t in (-1, 1); pyraft/sinogram/sino.py:175-212) and calls the
drop-in ``raft_radon_slantstack`` symbol of libraft, which runs on the GPU.

Only the slant-stack discretisation is provided: ``radon_method.BRESENHAM`` (the
reference's default) and the Fourier methods raise ``NotImplementedError``.
"""
import ctypes
import math

import numpy

from .raftypes import (RAFT_IMAGE, check_legacy, image, libraft, make_RAFT_IMAGE, make_RAFT_MATRIX,
                       nthreads)

__all__ = ["radon_method", "radon", "radon_gpu"]


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
    """Radon transform of a pyraft ``image`` (slant-stack) or of an ellipse description.

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
            raise NotImplementedError("radon: the Bresenham projector is not part of this build; "
                                      "pass method=radon_method.SLANTSTACK")
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
