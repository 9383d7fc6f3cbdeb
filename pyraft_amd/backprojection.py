"""Backprojection (pyraft/backprojection/backprojection.py).

``backprojection_plan`` / ``backprojection`` keep pyraft's plan-tuple protocol
(pyraft/backprojection/backprojection.py:35-213); SLANTSTACK and BRESENHAM plans are
``(method, shape, padding)`` and run the drop-in ``raft_backprojection_slantstack`` /
``raft_backprojection_bresenham`` symbols on the GPU.  (pyraft's BRESENHAM plan is an
empty list, which its ``backprojection`` cannot index; here it is a working plan.)
``radon_transpose`` defaults to the Bresenham transpose as in pyraft
(pyraft/backprojection/backprojection.py:220-233); ``method=SLANTSTACK`` gives the
exact matched transpose of the slant-stack Radon, which pyraft does not offer.
"""
import ctypes

import numpy

from .raftypes import check_legacy, image, libraft, make_RAFT_IMAGE, nthreads
from .sinogram import radon_method

__all__ = ["backprojection_plan", "backprojection", "radon_transpose", "backprojection_gpu"]

_SINO_TL, _SINO_BR = (0.0, 1.0), (numpy.pi, -1.0)


def _sino_corners(sino):
    tl = getattr(sino, 'top_left', None) or _SINO_TL
    br = getattr(sino, 'bottom_right', None) or _SINO_BR
    return tuple(tl), tuple(br)


def _float64(arr):
    return numpy.asarray(arr, dtype=numpy.float64)


def backprojection_plan(sino_shape=(256, 180), shape=(256, 256), padding=2, method=radon_method.BST):
    """Plan tuple for ``backprojection`` (pyraft/backprojection/backprojection.py:35-166)."""
    if method in (radon_method.SLANTSTACK, radon_method.BRESENHAM):
        return (method, shape, padding, )
    if method in (radon_method.BST, radon_method.ANDERSSON, radon_method.ANDERSSONPB):
        raise NotImplementedError("backprojection_plan: only the SLANTSTACK and BRESENHAM methods are "
                                  "implemented in pyraft_amd")
    raise TypeError('Unsupported method for Backprojection!')


def backprojection(sino, plan, **kwargs):
    """Backprojection into an ``image`` of ``plan[1]`` (corners from kwargs): slant-stack,
    or the line-traversal one normalised by R^T 1 (0/0 where no ray passes a pixel)."""
    method, shape = plan[0], plan[1]
    if method == radon_method.SLANTSTACK:
        fn, name = libraft.raft_backprojection_slantstack, "raft_backprojection_slantstack"
    elif method == radon_method.BRESENHAM:
        fn, name = libraft.raft_backprojection_bresenham, "raft_backprojection_bresenham"
    else:
        raise TypeError('Unsupported method for Backprojection!')
    img = image(shape, **kwargs)
    IMG = make_RAFT_IMAGE(img, img.top_left, img.bottom_right)
    tl, br = _sino_corners(sino)
    s64 = _float64(sino)
    SINO = make_RAFT_IMAGE(s64, tl, br)
    fn(SINO, IMG, ctypes.c_int(nthreads))
    check_legacy(name)
    return img


def radon_transpose(sino, shape=(256, 256), method=radon_method.BRESENHAM, **kwargs):
    """Transpose of ``radon`` for the method (adjoint to rounding)."""
    if method == radon_method.BRESENHAM:
        fn, name = libraft.raft_radon_transpose_bresenham, "raft_radon_transpose_bresenham"
    elif method == radon_method.SLANTSTACK:
        fn, name = libraft.raft_radon_transpose_slantstack, "raft_radon_transpose_slantstack"
    else:
        raise TypeError('Unsupported method for transpose Radon!')
    img = image(shape, **kwargs)
    IMG = make_RAFT_IMAGE(img, img.top_left, img.bottom_right)
    tl, br = _sino_corners(sino)
    s64 = _float64(sino)
    fn(make_RAFT_IMAGE(s64, tl, br), IMG, ctypes.c_int(nthreads))
    check_legacy(name)
    return img


def backprojection_gpu(sino, shape=(256, 256), **kwargs):
    """backprojection_gpu(sino, shape) (pyraft/backprojection/backprojection.py:241-261):
    float32 host round trip with the default geometry; returns a float64 ndarray."""
    img_size = shape[0]
    nrays, nviews = sino.shape[0], sino.shape[1]
    sino_buff = numpy.ascontiguousarray(sino, dtype=numpy.float32)
    img = numpy.zeros(img_size * img_size, dtype=numpy.float32)
    libraft.raft_backprojection_slantstack_gpu(img.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                               sino_buff.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                               img_size, nrays, nviews)
    check_legacy("raft_backprojection_slantstack_gpu")
    return img.reshape([img_size, img_size]).astype('float64')
