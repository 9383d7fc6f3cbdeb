"""FP64 CPU oracle for the PyRAFT slant-stack path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker or the timed CPU
baseline.  The shipped package ``pyraft_amd`` never imports it.

Contents (each restates the cited reference routine):

* ``radon`` / ``backprojection`` / ``radon_transpose`` -- ctypes calls into
  ``liboracle.so`` (oracle.c), restating raft/raft_radon_slantstack.c:122-214 and
  raft/raft_backprojection_slantstack.c:121-181, plus the exact adjoint of the
  forward (no reference counterpart).  ``backprojection_rows`` evaluates the same
  backprojection on a list of image rows only (full-size checks on a row sample).
* ``lowpassino_fbp`` / ``lowpassino_reg_fbp`` -- numpy restatements of
  pyraft/filters/filters.py:36-68 and 73-111.
* ``filter1d_ramp`` -- numpy restatement of raft/raft_filter.cpp:118-213 (the reference
  filters through FFTW, which this image lacks: parity unpinned beyond the FFT
  conventions, which tests/test_filters.py checks against a brute-force DFT).
* ``fbp`` -- filter then backprojection, the composition pyraft/inversion/inversion.py:48-56
  performs (with the CPU slant-stack backprojection in place of the CUDA one).
* ``em`` / ``osem`` -- the multiplicative EM recurrence of raft/raft_em_gpu.cu:41-69
  (x0 = 1, W = BP(1), x <- x * BP(s / R x) / W) composed from the slant-stack
  Radon and backprojection; OSEM walks contiguous angle blocks each with its own
  weight BP_b(1) (pyraft/iterative/em.py:48-57, 282-318).  Zero denominators use
  the convention documented below (the reference divides unguarded,
  raft/raft_cuda_aux.cu:21).
* ``loglik`` -- emission log-likelihood sum s*log(p) - sum p (xfct/likelihood.c:83-94).
* ``phantom_fromdesc`` / ``radon_fromdesc`` -- raft/raft_image_phantom_fromdesc.c:4-59,
  raft/raft_radon_fromdesc.c:4-67.
* ``radon_bresenham`` / ``radon_transpose_bresenham`` / ``backprojection_bresenham`` --
  the line-traversal pair, raft/raft_radon_bresenham.cpp, raft/raft_radon_transpose_bresenham.cpp,
  raft/raft_backprojection_bresenham.cpp over raft/raft_ddaiterator.c.
* ``radon_withmesh`` / ``radon_view`` -- the arbitrary-mesh forms,
  raft/raft_radon_slantstack_withmesh.c:49-224 and raft/raft_radon_slantstack_view.c:22-113.
* ``ref_*`` -- the same calls into ``_ref/libraft_ref.so``, the reference's own C
  compiled from its sources by ``oracle/Makefile`` (present only where it was built).

Conventions shared with the product (stated, not taken from the product):
  EM ratio   rho = s / p  where p > 0, else 0.
  EM update  x <- x * BP_b(rho) / W_b  where W_b > 0; pixels with W_b == 0 keep x.
  LL         sum over rays with p > 0 of s*log(p) - p, accumulated in FP64.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


class RaftMatrix(ctypes.Structure):
    _fields_ = [("p_data", ctypes.POINTER(ctypes.c_double)), ("lines", ctypes.c_int),
                ("line_stride", ctypes.c_int), ("columns", ctypes.c_int),
                ("column_stride", ctypes.c_int)]


class RaftImage(ctypes.Structure):
    _fields_ = [("data", RaftMatrix), ("tl_x", ctypes.c_double), ("tl_y", ctypes.c_double),
                ("br_x", ctypes.c_double), ("br_y", ctypes.c_double)]


class RaftVector(ctypes.Structure):
    _fields_ = [("p_data", ctypes.POINTER(ctypes.c_double)), ("size", ctypes.c_int), ("stride", ctypes.c_int)]


def _vector(v):
    assert v.dtype == np.float64 and v.ndim == 1
    return RaftVector(v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), v.shape[0], v.strides[0] // 8)


def _matrix(a):
    assert a.dtype == np.float64 and a.ndim == 2
    return RaftMatrix(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0],
                      a.strides[1] // 8, a.shape[1], a.strides[0] // 8)


def _image(a, tl, br):
    return RaftImage(_matrix(a), float(tl[0]), float(tl[1]), float(br[0]), float(br[1]))


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        _lib = ctypes.CDLL(path)
        for name in ("oracle_radon_slantstack", "oracle_backprojection_slantstack",
                     "oracle_radon_transpose_slantstack"):
            f = getattr(_lib, name)
            f.argtypes = [RaftImage, RaftImage, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            f.restype = None
        for name in ("oracle_phantom_fromdesc", "oracle_radon_fromdesc"):
            f = getattr(_lib, name)
            f.argtypes = [RaftImage, RaftMatrix]
            f.restype = None
        _lib.oracle_radon_withmesh.argtypes = [RaftImage, RaftImage, RaftVector, RaftVector, ctypes.c_int]
        _lib.oracle_radon_withmesh.restype = None
        _lib.oracle_radon_view.argtypes = [RaftVector, RaftImage, RaftVector, ctypes.c_double]
        _lib.oracle_radon_view.restype = None
        for name in ("oracle_radon_bresenham",):
            getattr(_lib, name).argtypes = [RaftImage, RaftImage, ctypes.c_int]
            getattr(_lib, name).restype = None
        for name in ("oracle_radon_transpose_bresenham", "oracle_backprojection_bresenham"):
            getattr(_lib, name).argtypes = [RaftImage, RaftImage, ctypes.c_int]
            getattr(_lib, name).restype = None
        _lib.oracle_dda_ray_path.argtypes = [RaftImage, RaftImage, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_int]
        _lib.oracle_dda_ray_path.restype = ctypes.c_int
        _lib.oracle_radon_sample_map_mesh.argtypes = [RaftImage, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        _lib.oracle_radon_sample_map_mesh.restype = None
        _lib.oracle_radon_sample_map.argtypes = [RaftImage, RaftImage, ctypes.c_void_p, ctypes.c_void_p]
        _lib.oracle_radon_sample_map.restype = None
        _lib.oracle_bp_sample_map.argtypes = [RaftImage, RaftImage, ctypes.c_void_p]
        _lib.oracle_bp_sample_map.restype = None
        _lib.oracle_backprojection_rows.argtypes = [RaftImage, RaftImage, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
        _lib.oracle_backprojection_rows.restype = None
    return _lib


def ref_path():
    return os.path.join(_HERE, "_ref", "libraft_ref.so")


def ref_available():
    return os.path.exists(ref_path())


def ref():
    """The reference's own C (compiled by oracle/Makefile), or None."""
    global _ref
    if _ref is None and ref_available():
        _ref = ctypes.CDLL(ref_path())
        for name in ("raft_radon_slantstack", "raft_backprojection_slantstack"):
            f = getattr(_ref, name)
            f.argtypes = [RaftImage, RaftImage, ctypes.c_int]
            f.restype = None
        for name in ("raft_image_phantom_fromdesc", "raft_radon_fromdesc"):
            f = getattr(_ref, name)
            f.argtypes = [RaftImage, RaftMatrix]
            f.restype = None
        # (phantom, radon, t, theta, nthreads): the definition's order, which pyraft uses
        _ref.raft_radon_slantstack_withmesh.argtypes = [RaftImage, RaftImage, RaftVector, RaftVector, ctypes.c_int]
        _ref.raft_radon_slantstack_withmesh.restype = None
        _ref.raft_radon_slantstack_view.argtypes = [RaftVector, RaftImage, RaftVector, ctypes.c_double]
        _ref.raft_radon_slantstack_view.restype = None
        _ref.harness_image_write.argtypes = [ctypes.c_char_p, RaftImage]
        _ref.harness_image_write.restype = ctypes.c_int
        _ref.harness_image_read.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_long, ctypes.c_void_p,
                                            ctypes.c_void_p]
        _ref.harness_image_read.restype = ctypes.c_int
        _ref.harness_radon_bresenham.argtypes = [RaftImage, RaftImage]
        _ref.harness_radon_bresenham.restype = None
        _ref.harness_radon_transpose_bresenham.argtypes = [RaftImage, RaftImage]
        _ref.harness_radon_transpose_bresenham.restype = None
        _ref.harness_backprojection_bresenham.argtypes = [RaftImage, RaftImage, ctypes.c_void_p]
        _ref.harness_backprojection_bresenham.restype = None
    return _ref


IMG_TL, IMG_BR = (-1.0, 1.0), (1.0, -1.0)
SINO_TL, SINO_BR = (0.0, 1.0), (np.pi, -1.0)


def _threads(nthreads):
    """Default: one thread per core, at most 16 (a GPU host's share per GPU)."""
    return int(nthreads) if nthreads else min(16, os.cpu_count() or 1)


# ---------------------------------------------------------------- projectors

def radon(img, nrays, nangles, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=SINO_TL, sino_br=SINO_BR,
          j0=0, j1=None, nthreads=None, out=None):
    """Slant-stack Radon (raft/raft_radon_slantstack.c:44-214); returns (nrays, nangles) FP64."""
    img = np.ascontiguousarray(img, dtype=np.float64)
    sino = np.zeros((nrays, nangles)) if out is None else out
    j1 = nangles if j1 is None else j1
    lib().oracle_radon_slantstack(_image(img, img_tl, img_br), _image(sino, sino_tl, sino_br),
                                  j0, j1, _threads(nthreads))
    return sino


def backprojection(sino, n, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=SINO_TL, sino_br=SINO_BR,
                   j0=0, j1=None, nthreads=None):
    """Slant-stack backprojection (raft/raft_backprojection_slantstack.c:43-181)."""
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    img = np.zeros((n, n))
    j1 = sino.shape[1] if j1 is None else j1
    lib().oracle_backprojection_slantstack(_image(sino, sino_tl, sino_br), _image(img, img_tl, img_br),
                                           j0, j1, _threads(nthreads))
    return img


def backprojection_rows(sino, n, rows, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=SINO_TL, sino_br=SINO_BR,
                       j0=0, j1=None, nthreads=None):
    """Rows `rows` of ``backprojection(sino, n, ...)``: a (len(rows), n) FP64 array,
    each pixel computed exactly as in the full image."""
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int32).reshape(-1)
    if rows.size and (rows.min() < 0 or rows.max() >= n):
        raise ValueError("rows outside [0, %d)" % n)
    geom = np.zeros((n, 1))  # corners and the line count n; the data are not read
    geom_img = RaftImage(RaftMatrix(geom.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n, 1, 1, 1),
                         float(img_tl[0]), float(img_tl[1]), float(img_br[0]), float(img_br[1]))
    out = np.zeros((rows.size, n))
    j1 = sino.shape[1] if j1 is None else j1
    lib().oracle_backprojection_rows(_image(sino, sino_tl, sino_br), geom_img, j0, j1, rows.ctypes.data,
                                     int(rows.size), out.ctypes.data, _threads(nthreads))
    return out


def radon_transpose(sino, n, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=SINO_TL, sino_br=SINO_BR,
                    j0=0, j1=None, nthreads=None):
    """Exact adjoint of ``radon`` (same index pattern, scatter form)."""
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    img = np.zeros((n, n))
    j1 = sino.shape[1] if j1 is None else j1
    lib().oracle_radon_transpose_slantstack(_image(sino, sino_tl, sino_br), _image(img, img_tl, img_br),
                                            j0, j1, _threads(nthreads))
    return img


def radon_sample_map(n, nrays, nangles, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=SINO_TL, sino_br=SINO_BR):
    """(mask, coord) of shape (A, R, N): which samples the reference accumulates and
    its FP64 interpolation coordinate for each."""
    img = np.zeros((n, n))
    sino = np.zeros((nrays, nangles))
    mask = np.zeros((nangles, nrays, n), dtype=np.uint8)
    coord = np.zeros((nangles, nrays, n))
    lib().oracle_radon_sample_map(_image(img, img_tl, img_br), _image(sino, sino_tl, sino_br),
                                  mask.ctypes.data, coord.ctypes.data)
    return mask.astype(bool), coord


def radon_sample_map_mesh(n, t, theta, img_tl=IMG_TL, img_br=IMG_BR):
    """radon_sample_map for rays at t[k] and angles theta[j] from tables."""
    t = np.ascontiguousarray(t, dtype=np.float64)
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    R, A = t.shape[0], theta.shape[0]
    mask = np.zeros((A, R, n), np.uint8)
    coord = np.zeros((A, R, n))
    dummy = np.zeros((n, n))
    lib().oracle_radon_sample_map_mesh(_image(dummy, img_tl, img_br), R, t.ctypes.data, A, theta.ctypes.data,
                                       mask.ctypes.data, coord.ctypes.data)
    return mask.astype(bool), coord


def bp_sample_map(n, nrays, nangles, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=SINO_TL, sino_br=SINO_BR):
    """mask of shape (A, N, N): which (angle, pixel) pairs the reference backprojection uses."""
    img = np.zeros((n, n))
    sino = np.zeros((nrays, nangles))
    mask = np.zeros((nangles, n, n), dtype=np.uint8)
    lib().oracle_bp_sample_map(_image(sino, sino_tl, sino_br), _image(img, img_tl, img_br), mask.ctypes.data)
    return mask.astype(bool)


# ---------------------------------------------------------------- phantoms

def phantom_fromdesc(desc, n, tl=IMG_TL, br=IMG_BR):
    desc = np.ascontiguousarray(desc, dtype=np.float64)
    img = np.zeros((n, n))
    lib().oracle_phantom_fromdesc(_image(img, tl, br), _matrix(desc))
    return img


def radon_fromdesc(desc, nrays, nangles, tl=SINO_TL, br=SINO_BR):
    desc = np.ascontiguousarray(desc, dtype=np.float64)
    sino = np.zeros((nrays, nangles))
    lib().oracle_radon_fromdesc(_image(sino, tl, br), _matrix(desc))
    return sino


# ---------------------------------------------------------------- reference C

def radon_withmesh(img, t, theta, img_tl=IMG_TL, img_br=IMG_BR, nthreads=None):
    """Slant-stack Radon on arbitrary tables (raft/raft_radon_slantstack_withmesh.c:49-224):
    element (k, j) = ray sum at theta[j], t[k]; returns (len(t), len(theta)) FP64."""
    img = np.ascontiguousarray(img, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    sino = np.zeros((t.shape[0], theta.shape[0]))
    lib().oracle_radon_withmesh(_image(img, img_tl, img_br), _image(sino, (0.0, 1.0), (1.0, -1.0)), _vector(t),
                                _vector(theta), _threads(nthreads))
    return sino


def radon_view(img, t, theta, img_tl=IMG_TL, img_br=IMG_BR):
    """One angle (raft/raft_radon_slantstack_view.c:22-113); returns len(t) FP64."""
    img = np.ascontiguousarray(img, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    view = np.zeros(t.shape[0])
    lib().oracle_radon_view(_vector(view), _image(img, img_tl, img_br), _vector(t), float(theta))
    return view


def ref_radon_withmesh(img, t, theta, img_tl=IMG_TL, img_br=IMG_BR, nthreads=1):
    """The reference's withmesh; nthreads must divide len(t)*len(theta) for every output
    to be written (its thread split drops the remainder)."""
    img = np.ascontiguousarray(img, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    sino = np.zeros((t.shape[0], theta.shape[0]))
    ref().raft_radon_slantstack_withmesh(_image(img, img_tl, img_br), _image(sino, (0.0, 1.0), (1.0, -1.0)),
                                         _vector(t), _vector(theta), nthreads)
    return sino


def ref_radon_view(img, t, theta, img_tl=IMG_TL, img_br=IMG_BR):
    img = np.ascontiguousarray(img, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    view = np.zeros(t.shape[0])
    ref().raft_radon_slantstack_view(_vector(view), _image(img, img_tl, img_br), _vector(t), float(theta))
    return view


# Line-traversal (Bresenham / DDA) pair.  Sinogram corners as pyraft.radon sets them by
# default (theta_extent (0, pi), t from +1 at row 0 to -1, pyraft/sinogram/sino.py:179-193);
# the angle pitch divides by A - 1 here (raft/raft_image.h:56-61).
DDA_SINO_TL, DDA_SINO_BR = (0.0, 1.0), (np.pi, -1.0)


def radon_bresenham(img, nrays, nangles, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=DDA_SINO_TL,
                    sino_br=DDA_SINO_BR, nthreads=None):
    """raft/raft_radon_bresenham.cpp:60-105 (traversal raft/raft_ddaiterator.c)."""
    img = np.ascontiguousarray(img, dtype=np.float64)
    sino = np.zeros((nrays, nangles))
    lib().oracle_radon_bresenham(_image(img, img_tl, img_br), _image(sino, sino_tl, sino_br), _threads(nthreads))
    return sino


def radon_transpose_bresenham(sino, shape, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=DDA_SINO_TL,
                              sino_br=DDA_SINO_BR, nthreads=None, out=None):
    """raft/raft_radon_transpose_bresenham.cpp: out += R^T sino (out defaults to zeros)."""
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    img = np.zeros(shape) if out is None else out
    lib().oracle_radon_transpose_bresenham(_image(sino, sino_tl, sino_br), _image(img, img_tl, img_br),
                                           _threads(nthreads))
    return img


def backprojection_bresenham(sino, shape, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=DDA_SINO_TL,
                             sino_br=DDA_SINO_BR, nthreads=None):
    """raft/raft_backprojection_bresenham.cpp:120-251: R^T sino / R^T 1 (0/0 where no ray passes)."""
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    img = np.zeros(shape)
    with np.errstate(all="ignore"):
        lib().oracle_backprojection_bresenham(_image(sino, sino_tl, sino_br), _image(img, img_tl, img_br),
                                              _threads(nthreads))
    return img


def dda_ray_path(shape, nrays, nangles, a, r, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=DDA_SINO_TL,
                 sino_br=DDA_SINO_BR):
    """(pixels (k, 2) as (line, column), lengths (k,)) of ray r of angle a, in visit order."""
    cap = 2 * (shape[0] + shape[1]) + 8
    ij = np.zeros((cap, 2), np.int32)
    ln = np.zeros(cap)
    dummy_img = np.zeros((2, 2))
    dummy_sino = np.zeros((2, 2))
    img = _image(dummy_img, img_tl, img_br)
    img.data.lines, img.data.columns = shape[0], shape[1]
    sino = _image(dummy_sino, sino_tl, sino_br)
    sino.data.lines, sino.data.columns = nrays, nangles
    k = lib().oracle_dda_ray_path(img, sino, a, r, ij.ctypes.data, ln.ctypes.data, cap)
    return ij[:k].copy(), ln[:k].copy()


def ref_radon_bresenham(img, nrays, nangles, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=DDA_SINO_TL,
                        sino_br=DDA_SINO_BR):
    img = np.ascontiguousarray(img, dtype=np.float64)
    sino = np.zeros((nrays, nangles))
    ref().harness_radon_bresenham(_image(img, img_tl, img_br), _image(sino, sino_tl, sino_br))
    return sino


def ref_radon_transpose_bresenham(sino, shape, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=DDA_SINO_TL,
                                  sino_br=DDA_SINO_BR):
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    img = np.zeros(shape)
    ref().harness_radon_transpose_bresenham(_image(sino, sino_tl, sino_br), _image(img, img_tl, img_br))
    return img


def ref_backprojection_bresenham(sino, shape, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=DDA_SINO_TL,
                                 sino_br=DDA_SINO_BR):
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    img = np.zeros(shape)
    scale = np.zeros(shape)
    with np.errstate(all="ignore"):
        ref().harness_backprojection_bresenham(_image(sino, sino_tl, sino_br), _image(img, img_tl, img_br),
                                               scale.ctypes.data)
    return img


def ref_image_write(path, img, tl, br):
    """The reference's raft_image_write (raft/raft_image.c:135-151) to a file."""
    img = np.ascontiguousarray(img, dtype=np.float64)
    return ref().harness_image_write(str(path).encode(), _image(img, tl, br))


def ref_image_read(path, cap=1 << 22):
    """The reference's raft_image_read (raft/raft_image.c:5-38): (values, corners) or None."""
    out = np.zeros(cap)
    shape = np.zeros(2, np.int32)
    corners = np.zeros(4)
    rc = ref().harness_image_read(str(path).encode(), out.ctypes.data, cap, shape.ctypes.data, corners.ctypes.data)
    if rc:
        return None
    return out[:shape[0] * shape[1]].reshape(shape[0], shape[1]).copy(), tuple(corners)


def ref_radon(img, nrays, nangles, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=SINO_TL, sino_br=SINO_BR,
              nthreads=1):
    img = np.ascontiguousarray(img, dtype=np.float64)
    sino = np.zeros((nrays, nangles))
    ref().raft_radon_slantstack(_image(img, img_tl, img_br), _image(sino, sino_tl, sino_br), nthreads)
    return sino


def ref_backprojection(sino, n, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=SINO_TL, sino_br=SINO_BR,
                       nthreads=1):
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    img = np.zeros((n, n))
    ref().raft_backprojection_slantstack(_image(sino, sino_tl, sino_br), _image(img, img_tl, img_br),
                                         nthreads)
    return img


def ref_phantom_fromdesc(desc, n, tl=IMG_TL, br=IMG_BR):
    desc = np.ascontiguousarray(desc, dtype=np.float64)
    img = np.zeros((n, n))
    ref().raft_image_phantom_fromdesc(_image(img, tl, br), _matrix(desc))
    return img


def ref_radon_fromdesc(desc, nrays, nangles, tl=SINO_TL, br=SINO_BR):
    desc = np.ascontiguousarray(desc, dtype=np.float64)
    sino = np.zeros((nrays, nangles))
    ref().raft_radon_fromdesc(_image(sino, tl, br), _matrix(desc))
    return sino


# ---------------------------------------------------------------- FBP

def lowpassino_fbp(sino, tl=SINO_TL, br=SINO_BR, t=None):
    """pyraft/filters/filters.py:36-68: per column Re(ifft(fft(s) * fftshift(h)))."""
    sino = np.asarray(sino, dtype=np.float64)
    tmin, tmax = br[1], tl[1]
    if t is None:
        t = np.linspace(tmin, tmax, sino.shape[0])
    R, V = int(sino.shape[0]), int(sino.shape[1])
    dt = (t[R - 1] - t[0]) / (R - 1)
    wc = 1 / (2 * dt)
    w = np.linspace(-wc, wc, R)
    h = np.abs(w) * np.cos(np.pi * w * dt)
    G = np.fft.fftshift(np.transpose(np.kron(np.ones((V, 1)), h)))
    return np.fft.ifft(np.fft.fft(sino, axis=0) * G, axis=0).real


def ramp_kernel(R, tl=SINO_TL, br=SINO_BR):
    """The real circular-convolution kernel g = Re(ifft(fftshift(h))) equivalent to
    ``lowpassino_fbp`` (s is real, so only the even part of the spectrum acts)."""
    tmin, tmax = br[1], tl[1]
    t = np.linspace(tmin, tmax, R)
    dt = (t[R - 1] - t[0]) / (R - 1)
    wc = 1 / (2 * dt)
    w = np.linspace(-wc, wc, R)
    h = np.abs(w) * np.cos(np.pi * w * dt)
    return np.fft.ifft(np.fft.fftshift(h)).real


def lowpassino_reg_fbp(sino, L, tl=SINO_TL, br=SINO_BR):
    """pyraft/filters/filters.py:73-111: as lowpassino_fbp with h = |w| / (1 + L |w|^2)."""
    sino = np.asarray(sino, dtype=np.float64)
    tmin, tmax = br[1], tl[1]
    t = np.linspace(tmin, tmax, sino.shape[0])
    R, V = int(sino.shape[0]), int(sino.shape[1])
    dt = (t[R - 1] - t[0]) / (R - 1)
    wc = 1 / (2 * dt)
    w = np.linspace(-wc, wc, R)
    h = np.abs(w) / (1.0 + L * np.abs(w) ** 2.0)
    G = np.fft.fftshift(np.transpose(np.kron(np.ones((V, 1)), h)))
    return np.fft.ifft(np.fft.fft(sino, axis=0) * G, axis=0).real


def filter1d_ramp_vector(R, cutoff, tl=SINO_TL, br=SINO_BR):
    """The ramp of raft/raft_filter.cpp:184-206: F = R/2 + 1 values, delta * i for
    i < int(F * min(cutoff, 1)) and 0 after, delta = pi / (F - 1) / |dt| with
    dt = (br_y - tl_y) / (R - 1) (raft/raft_image.h:60-61)."""
    F = R // 2 + 1
    cutoff = cutoff if cutoff <= 1.0 else 1.0
    delta = 3.14159265358979323846264338327 / (F - 1)
    delta /= abs((br[1] - tl[1]) / (R - 1))
    length = int(F * cutoff)
    filt = np.zeros(F)
    if length > 0:
        filt[:min(length, F)] = delta * np.arange(min(length, F))
    return filt


def filter1d_ramp(sino, cutoff, tl=SINO_TL, br=SINO_BR):
    """raft/raft_filter.cpp:118-213 restated with numpy's FFT (FFTW is absent here): per
    column, zero-pad to Z = 2R - 1, real FFT, multiply bins 0..R/2 by ramp/Z (bins
    R/2+1..R-1 untouched, as the reference's loop bound osize = R/2 + 1 leaves them),
    unnormalised inverse (FFTW's c2r: numpy's irfft times Z), keep the first R samples."""
    sino = np.asarray(sino, dtype=np.float64)
    R = sino.shape[0]
    F = R // 2 + 1
    filt = filter1d_ramp_vector(R, cutoff, tl, br)
    Z = 2 * R - 1
    xz = np.zeros((Z, sino.shape[1]))
    xz[:R] = sino
    X = np.fft.rfft(xz, axis=0)
    X[:F] *= (filt / Z)[:, None]
    return (np.fft.irfft(X, n=Z, axis=0) * Z)[:R]


def fbp(sino, n, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=SINO_TL, sino_br=SINO_BR, nthreads=None):
    filt = lowpassino_fbp(sino, sino_tl, sino_br)
    return backprojection(filt, n, img_tl, img_br, sino_tl, sino_br, nthreads=nthreads)


# ---------------------------------------------------------------- transmission EM

def tr_em(raw, flat, n, niter, x0=None, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=SINO_TL, sino_br=SINO_BR):
    """raft/raft_transmission_em.cu:16-66 on the slant-stack pair: W = BP(raw), then
    x <- (x r) r with r = BP(flat exp(-R x)) / W (the update applied twice, as the
    reference does); pixels with W = 0 keep their value."""
    R, A = raw.shape
    geo = dict(img_tl=img_tl, img_br=img_br, sino_tl=sino_tl, sino_br=sino_br)
    x = np.ones((n, n)) if x0 is None else np.array(x0, dtype=np.float64)
    w = backprojection(np.asarray(raw, np.float64), n, **geo)
    ok = w != 0
    for _ in range(niter):
        p = np.asarray(flat, np.float64) * np.exp(-radon(x, R, A, **geo))
        b = backprojection(p, n, **geo)
        r = np.where(ok, b / np.where(ok, w, 1.0), 1.0)
        x = (x * r) * r
    return x


# ---------------------------------------------------------------- EM / OSEM

def subset_bounds(nangles, nsubsets):
    """Contiguous angle blocks [b*A//S, (b+1)*A//S) (pyraft/iterative/em.py:284-312)."""
    return [((b * nangles) // nsubsets, ((b + 1) * nangles) // nsubsets) for b in range(nsubsets)]


def loglik(s, p):
    """xfct/likelihood.c:83-94, over rays with p > 0."""
    s = np.asarray(s, dtype=np.float64)
    p = np.asarray(p, dtype=np.float64)
    m = p > 0
    return float(np.sum(s[m] * np.log(p[m])) - np.sum(p[m]))


def osem(sino, n, niter, nsubsets=1, x0=None, img_tl=IMG_TL, img_br=IMG_BR, sino_tl=SINO_TL,
         sino_br=SINO_BR, nthreads=None, track_ll=True):
    """Ordered-subsets EM with the raft_em_gpu recurrence (nsubsets=1 is plain EM).

    Returns (x, ll) where ll[k] sums the subset log-likelihood terms seen during
    iteration k (for nsubsets=1 this is LL(x_k) before the update)."""
    sino = np.ascontiguousarray(sino, dtype=np.float64)
    R, A = sino.shape
    x = np.ones((n, n)) if x0 is None else np.array(x0, dtype=np.float64)
    geo = dict(img_tl=img_tl, img_br=img_br, sino_tl=sino_tl, sino_br=sino_br, nthreads=nthreads)
    blocks = subset_bounds(A, nsubsets)
    ones = np.ones_like(sino)
    weights = [backprojection(ones, n, j0=b0, j1=b1, **geo) for (b0, b1) in blocks]
    ll = []
    for _ in range(niter):
        acc = 0.0
        for (b0, b1), W in zip(blocks, weights):
            p = radon(x, R, A, j0=b0, j1=b1, **geo)[:, b0:b1]
            s = sino[:, b0:b1]
            if track_ll:
                acc += loglik(s, p)
            rho = np.zeros((R, A))
            rho[:, b0:b1] = np.where(p > 0, s / np.where(p > 0, p, 1.0), 0.0)
            bp = backprojection(rho, n, j0=b0, j1=b1, **geo)
            x = np.where(W > 0, x * bp / np.where(W > 0, W, 1.0), x)
        ll.append(acc)
    return x, ll


def em(sino, n, niter, **kw):
    return osem(sino, n, niter, nsubsets=1, **kw)
