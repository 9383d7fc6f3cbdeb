"""Device-resident API over the libraft ``raft_ss_*`` entry points.

PyTorch supplies device memory and the current stream; every operator runs a
hand-written HIP kernel of libraft through ctypes (``data_ptr()`` in, status code
out).  Tensors are float32 and contiguous:

    image     (..., N, N)   row r <-> y = ymin + r*dy (row 0 at the bottom)
    sinogram  (..., R, A)   the pyraft (rays, angles) layout
    sinogramT (..., A, R)   angle-major (``angle_major=True`` variants)

Leading dimensions are flattened into the slice count of one launch.
"""
import ctypes
import math

import torch

from .raftypes import BP_NVARIANTS, TUNE_FORCE, check, libraft

DEFAULT_IMG_CORNERS = (-1.0, 1.0, 1.0, -1.0)
DEFAULT_SINO_CORNERS = (0.0, 1.0, math.pi, -1.0)


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def current_stream_handle(device=None):
    """hipStream_t of torch's current stream on `device` (default: the current device)."""
    if _raw_stream is not None:
        idx = None
        if device is not None:
            idx = torch.device(device).index if not isinstance(device, int) else device
        return ctypes.c_void_p(_raw_stream(torch.cuda.current_device() if idx is None else idx))
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class _on_device:
    """torch.cuda.device(dev) that skips the switch when dev is already current."""
    __slots__ = ("index", "prev")

    def __init__(self, dev):
        self.index = dev.index

    def __enter__(self):
        self.prev = torch.cuda.current_device()
        if self.prev != self.index:
            torch.cuda.set_device(self.index)

    def __exit__(self, *exc):
        if self.prev != self.index:
            torch.cuda.set_device(self.prev)
        return False


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("pyraft_amd.device needs a ROCm GPU (torch.cuda.is_available() is False)")


def _check_tensor(t, shape_tail, name):
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch.Tensor" % name)
    if t.dtype != torch.float32:
        raise TypeError("%s must be float32 (got %s)" % (name, t.dtype))
    if t.device.type != "cuda":
        raise TypeError("%s must be a device tensor" % name)
    if tuple(t.shape[-len(shape_tail):]) != tuple(shape_tail):
        raise ValueError("%s has shape %s, expected (..., %s)" % (name, tuple(t.shape),
                                                                  ", ".join(map(str, shape_tail))))
    return t.contiguous()


def _check_out(t, shape, device, name):
    """An output tensor the kernels write through its data pointer: float32, on the
    input's device, contiguous and of exactly `shape` (a copy would drop the results)."""
    if not isinstance(t, torch.Tensor):
        raise TypeError("%s must be a torch.Tensor" % name)
    if t.dtype != torch.float32:
        raise TypeError("%s must be float32 (got %s)" % (name, t.dtype))
    if t.device != device:
        raise TypeError("%s is on %s, the input on %s" % (name, t.device, device))
    if tuple(t.shape) != tuple(shape):
        raise ValueError("%s has shape %s, expected %s" % (name, tuple(t.shape), tuple(shape)))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    return t


def _nslices(t, tail):
    n = 1
    for d in t.shape[:-tail]:
        n *= int(d)
    return n


class Plan:
    """Geometry plan of an N x N image and an R x A sinogram on one device.

    ``img_corners`` / ``sino_corners`` are (tl_x, tl_y, br_x, br_y) as in
    raft_image: the image spans x in [tl_x, br_x], y in [br_y, tl_y]; the sinogram
    spans theta = tl_x + j*(br_x - tl_x)/A and t = br_y + k*(tl_y - br_y)/(R - 1)."""

    def __init__(self, n, nrays, nangles, img_corners=DEFAULT_IMG_CORNERS,
                 sino_corners=DEFAULT_SINO_CORNERS, device=None):
        _require_gpu()
        self.n, self.nrays, self.nangles = int(n), int(nrays), int(nangles)
        self.img_corners = tuple(float(v) for v in img_corners)
        self.sino_corners = tuple(float(v) for v in sino_corners)
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else
                                   torch.device(device).index or 0)
        self._handle = ctypes.c_void_p()
        ic = (ctypes.c_double * 4)(*self.img_corners)
        sc = (ctypes.c_double * 4)(*self.sino_corners)
        with _on_device(self.device):
            check(libraft.raft_ss_plan_create(ctypes.byref(self._handle), self.n, self.nrays, self.nangles,
                                              ic, sc, current_stream_handle()), "raft_ss_plan_create")

    @classmethod
    def mesh(cls, n, t, theta, img_corners=DEFAULT_IMG_CORNERS, device=None):
        """Plan on arbitrary tables: ray k at detector coordinate ``t[k]``, angle j at
        ``theta[j]`` (the geometry of raft_radon_slantstack_withmesh,
        raft/raft_radon_slantstack_withmesh.c:176-224).  Only ``radon`` accepts it."""
        _require_gpu()
        import numpy
        ts = numpy.ascontiguousarray(t, dtype=numpy.float64).reshape(-1)
        ths = numpy.ascontiguousarray(theta, dtype=numpy.float64).reshape(-1)
        self = cls.__new__(cls)
        self.n, self.nrays, self.nangles = int(n), int(ts.shape[0]), int(ths.shape[0])
        self.img_corners = tuple(float(v) for v in img_corners)
        self.sino_corners = None
        self.t, self.theta = ts, ths
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else
                                   torch.device(device).index or 0)
        self._handle = ctypes.c_void_p()
        ic = (ctypes.c_double * 4)(*self.img_corners)
        dp = ctypes.POINTER(ctypes.c_double)
        with _on_device(self.device):
            check(libraft.raft_ss_plan_create_mesh(ctypes.byref(self._handle), self.n, ic, self.nrays,
                                                   ts.ctypes.data_as(dp), self.nangles, ths.ctypes.data_as(dp),
                                                   current_stream_handle()), "raft_ss_plan_create_mesh")
        return self

    @property
    def is_mesh(self):
        return self.sino_corners is None

    # -- bookkeeping
    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                libraft.raft_ss_plan_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._handle = None

    @property
    def handle(self):
        return self._handle

    def info(self):
        out = (ctypes.c_int * 9)()
        check(libraft.raft_ss_plan_info(self._handle, out, 9), "raft_ss_plan_info")
        keys = ("n", "nrays", "nangles", "irregular_rays", "degenerate_angles", "bp_window", "bp_angles",
                "bp_variant", "bp_groups")
        return dict(zip(keys, list(out)))

    def set_bp_variant(self, variant):
        """Use backprojection kernel variant ``variant`` (see ``bp_variant_shape``)."""
        check(libraft.raft_ss_plan_set_bp_variant(self._handle, int(variant)), "raft_ss_plan_set_bp_variant")

    def set_radon_slab_rows(self, rows):
        """Pin the sample rows per slab of this plan's divided Radon launches (rows >= n:
        no division; 0: back to the tuned choice)."""
        check(libraft.raft_ss_plan_set_radon_slab_rows(self._handle, int(rows)), "raft_ss_plan_set_radon_slab_rows")

    def tune(self, reps=3, force=False):
        """Time every backprojection variant on this geometry, keep the fastest.
        Returns (variant, {variant: microseconds}) (variants that do not fit are left out)."""
        times = (ctypes.c_double * BP_NVARIANTS)()
        with _on_device(self.device):
            v = libraft.raft_ss_plan_tune(self._handle, int(reps), TUNE_FORCE if force else 0, times, self._stream())
        check(v if v < 0 else 0, "raft_ss_plan_tune")
        return v, {i: times[i] for i in range(BP_NVARIANTS) if times[i] >= 0}

    def tune_transpose(self, nslices=1, reps=3, force=False):
        """Time the exact transpose's tile variants on `nslices` slices, keep the fastest.
        Returns (variant, {variant: microseconds}); variant -1 when the line form applies."""
        times = (ctypes.c_double * BP_NVARIANTS)()
        with _on_device(self.device):
            v = libraft.raft_ss_plan_tune_transpose(self._handle, int(nslices), int(reps),
                                                    TUNE_FORCE if force else 0, times, self._stream())
        if v < -1:
            check(v, "raft_ss_plan_tune_transpose")
        return v, {i: times[i] for i in range(BP_NVARIANTS) if times[i] >= 0}

    def bounds(self):
        import numpy
        b = numpy.zeros((self.nangles, self.nrays, 2), numpy.int32)
        check(libraft.raft_ss_plan_bounds(self._handle, b.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
              "raft_ss_plan_bounds")
        return b

    def wave_bounds(self):
        import numpy
        nw = (self.nrays + 63) // 64
        b = numpy.zeros((self.nangles, nw, 4), numpy.int32)
        check(libraft.raft_ss_plan_wave_bounds(self._handle, b.ctypes.data_as(ctypes.POINTER(ctypes.c_int))),
              "raft_ss_plan_wave_bounds")
        return b

    def _stream(self):
        return current_stream_handle(self.device)

    def _range(self, angles):
        if angles is None:
            return 0, self.nangles
        j0, j1 = int(angles[0]), int(angles[1])
        if not (0 <= j0 <= j1 <= self.nangles):
            raise ValueError("angle range (%d, %d) outside [0, %d]" % (j0, j1, self.nangles))
        return j0, j1

    # -- operators
    def radon(self, image, angle_major=False, angles=None, out=None):
        """Slant-stack Radon of (..., N, N) -> (..., R, A) (or (..., A, R))."""
        x = _check_tensor(image, (self.n, self.n), "image")
        lead = x.shape[:-2]
        shape = lead + ((self.nangles, self.nrays) if angle_major else (self.nrays, self.nangles))
        j0, j1 = self._range(angles)
        if out is None:
            # the kernels write every column of the range; the rest of a partial range is zero
            full = j0 == 0 and j1 == self.nangles
            out = (torch.empty if full else torch.zeros)(shape, dtype=torch.float32, device=x.device)
        else:
            _check_out(out, shape, x.device, "out")
        with _on_device(self.device):
            check(libraft.raft_ss_radon_range(self._handle, x.data_ptr(), out.data_ptr(), int(angle_major), j0, j1,
                                              _nslices(x, 2), self._stream()), "raft_ss_radon_range")
        return out

    def backprojection(self, sino, angle_major=False, angles=None):
        """Slant-stack backprojection (reference BP semantics) -> (..., N, N)."""
        tail = (self.nangles, self.nrays) if angle_major else (self.nrays, self.nangles)
        s = _check_tensor(sino, tail, "sinogram")
        out = torch.empty(s.shape[:-2] + (self.n, self.n), dtype=torch.float32, device=s.device)
        j0, j1 = self._range(angles)
        with _on_device(self.device):
            check(libraft.raft_ss_backprojection_range(self._handle, s.data_ptr(), out.data_ptr(), int(angle_major),
                                                       j0, j1, _nslices(s, 2), self._stream()),
                  "raft_ss_backprojection_range")
        return out

    def radon_transpose(self, sino, angle_major=False, angles=None):
        """Exact adjoint of ``radon`` -> (..., N, N)."""
        tail = (self.nangles, self.nrays) if angle_major else (self.nrays, self.nangles)
        s = _check_tensor(sino, tail, "sinogram")
        out = torch.empty(s.shape[:-2] + (self.n, self.n), dtype=torch.float32, device=s.device)
        j0, j1 = self._range(angles)
        with _on_device(self.device):
            check(libraft.raft_ss_radon_transpose_range(self._handle, s.data_ptr(), out.data_ptr(), int(angle_major),
                                                        j0, j1, _nslices(s, 2), self._stream()),
                  "raft_ss_radon_transpose_range")
        return out

    def fbp_filter(self, sino, angle_major=False):
        tail = (self.nangles, self.nrays) if angle_major else (self.nrays, self.nangles)
        s = _check_tensor(sino, tail, "sinogram")
        out = torch.empty_like(s)
        with _on_device(self.device):
            check(libraft.raft_ss_fbp_filter(self._handle, s.data_ptr(), out.data_ptr(), int(angle_major),
                                             _nslices(s, 2), self._stream()), "raft_ss_fbp_filter")
        return out

    def fbp(self, sino):
        s = _check_tensor(sino, (self.nrays, self.nangles), "sinogram")
        out = torch.empty(s.shape[:-2] + (self.n, self.n), dtype=torch.float32, device=s.device)
        with _on_device(self.device):
            check(libraft.raft_ss_fbp(self._handle, s.data_ptr(), out.data_ptr(), _nslices(s, 2), self._stream()),
                  "raft_ss_fbp")
        return out

    def em(self, sino, niter, nsubsets=1, x0=None, return_ll=False):
        """OSEM (nsubsets=1: EM) from x0 (default ones); returns x (and the LL trace)."""
        s = _check_tensor(sino, (self.nrays, self.nangles), "sinogram")
        lead = s.shape[:-2]
        if x0 is None:
            x = torch.ones(lead + (self.n, self.n), dtype=torch.float32, device=s.device)
        else:
            x = _check_tensor(x0, (self.n, self.n), "x0").clone()
        ll = (ctypes.c_double * max(1, int(niter)))()
        with _on_device(self.device):
            check(libraft.raft_ss_em(self._handle, s.data_ptr(), x.data_ptr(), int(niter), int(nsubsets),
                                     _nslices(s, 2), ll if return_ll else None, self._stream()), "raft_ss_em")
        if return_ll:
            return x, [ll[i] for i in range(int(niter))]
        return x


def bp_variant_shape(variant):
    """(pixel rows per thread, LDS stage pairs) of backprojection variant ``variant``;
    the block tile is 32 columns x 8*rows rows."""
    rows, stage = ctypes.c_int(), ctypes.c_int()
    check(libraft.raft_ss_bp_variant_shape(int(variant), ctypes.byref(rows), ctypes.byref(stage)),
          "raft_ss_bp_variant_shape")
    return rows.value, stage.value


def ramp_kernel(nrays, tmin=-1.0, tmax=1.0, kind=0, param=0.0):
    """FP32 circular-convolution kernel (length nrays) for t = linspace(tmin, tmax, nrays) of
    lowpassino_fbp (kind 0) or lowpassino_reg_fbp with L = param (kind 1)."""
    import numpy
    g = numpy.zeros(int(nrays), numpy.float32)
    check(libraft.raft_ss_filter_kernel(int(nrays), float(tmin), float(tmax), int(kind), float(param),
                                        g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))), "raft_ss_filter_kernel")
    return g


def filter_rows(rows, g):
    """Circular convolution of each contiguous row of ``rows`` (..., R) with ``g`` (R,)."""
    r = rows.contiguous()
    if r.dtype != torch.float32 or g.dtype != torch.float32:
        raise TypeError("filter_rows needs float32 tensors")
    R = r.shape[-1]
    out = torch.empty_like(r)
    check(libraft.raft_ss_filter_rows(r.data_ptr(), out.data_ptr(), g.contiguous().data_ptr(), int(R),
                                      int(r.numel() // R), current_stream_handle(r.device)), "raft_ss_filter_rows")
    return out


def phantom(desc, n, corners=DEFAULT_IMG_CORNERS, device=None):
    """Rasterise an ellipse description (K x 6) into a device float32 (n, n) image."""
    import numpy
    d = numpy.ascontiguousarray(desc, dtype=numpy.float64).reshape(-1, 6)
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    out = torch.empty((int(n), int(n)), dtype=torch.float32, device=dev)
    c = (ctypes.c_double * 4)(*[float(v) for v in corners])
    with torch.cuda.device(dev):
        check(libraft.raft_ss_phantom(d.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), d.shape[0], int(n), int(n),
                                      c, out.data_ptr(), current_stream_handle(dev)), "raft_ss_phantom")
    return out


class DdaPlan:
    """Line-traversal (Bresenham / DDA) geometry of a lines x cols image and an R x A
    sinogram on one device (raft/raft_ddaiterator.c, raft/raft_radon_bresenham.cpp).

    ``img_corners`` / ``sino_corners`` are (tl_x, tl_y, br_x, br_y): image line i at
    y = tl_y + i*(br_y - tl_y)/(lines - 1) (line 0 at the top for the default corners),
    angle a at theta = tl_x + a*(br_x - tl_x)/(A - 1), ray r at
    t = tl_y + r*(br_y - tl_y)/(R - 1)."""

    def __init__(self, shape, nrays, nangles, img_corners=DEFAULT_IMG_CORNERS,
                 sino_corners=DEFAULT_SINO_CORNERS, device=None):
        _require_gpu()
        if isinstance(shape, int):
            shape = (shape, shape)
        self.shape = (int(shape[0]), int(shape[1]))
        self.nrays, self.nangles = int(nrays), int(nangles)
        self.img_corners = tuple(float(v) for v in img_corners)
        self.sino_corners = tuple(float(v) for v in sino_corners)
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else
                                   torch.device(device).index or 0)
        self._handle = ctypes.c_void_p()
        ic = (ctypes.c_double * 4)(*self.img_corners)
        sc = (ctypes.c_double * 4)(*self.sino_corners)
        with _on_device(self.device):
            check(libraft.raft_dda_plan_create(ctypes.byref(self._handle), self.shape[0], self.shape[1], self.nrays,
                                               self.nangles, ic, sc, current_stream_handle()), "raft_dda_plan_create")

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                libraft.raft_dda_plan_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._handle = None

    def info(self):
        out = (ctypes.c_int * 7)()
        check(libraft.raft_dda_plan_info(self._handle, out, 7), "raft_dda_plan_info")
        return dict(zip(("lines", "cols", "nrays", "nangles", "line_fast_angles", "traced_angles", "straight_angles"),
                        list(out)))

    def _stream(self):
        return current_stream_handle(self.device)

    def radon(self, image):
        """(..., lines, cols) -> (..., R, A)."""
        x = _check_tensor(image, self.shape, "image")
        out = torch.empty(x.shape[:-2] + (self.nrays, self.nangles), dtype=torch.float32, device=x.device)
        with _on_device(self.device):
            check(libraft.raft_dda_radon(self._handle, x.data_ptr(), out.data_ptr(), _nslices(x, 2), self._stream()),
                  "raft_dda_radon")
        return out

    def radon_transpose(self, sino):
        """(..., R, A) -> (..., lines, cols), the transpose of ``radon``."""
        s = _check_tensor(sino, (self.nrays, self.nangles), "sinogram")
        out = torch.empty(s.shape[:-2] + self.shape, dtype=torch.float32, device=s.device)
        with _on_device(self.device):
            check(libraft.raft_dda_radon_transpose(self._handle, s.data_ptr(), out.data_ptr(), _nslices(s, 2),
                                                   self._stream()), "raft_dda_radon_transpose")
        return out

    def backprojection(self, sino):
        """(..., R, A) -> (..., lines, cols): R^T sino / R^T 1 (NaN where no ray passes)."""
        s = _check_tensor(sino, (self.nrays, self.nangles), "sinogram")
        out = torch.empty(s.shape[:-2] + self.shape, dtype=torch.float32, device=s.device)
        with _on_device(self.device):
            check(libraft.raft_dda_backprojection(self._handle, s.data_ptr(), out.data_ptr(), _nslices(s, 2),
                                                  self._stream()), "raft_dda_backprojection")
        return out

    def weights(self):
        """R^T 1, the backprojection's normalisation, (lines, cols)."""
        out = torch.empty(self.shape, dtype=torch.float32, device=self.device)
        with _on_device(self.device):
            check(libraft.raft_dda_weights(self._handle, out.data_ptr(), self._stream()), "raft_dda_weights")
        return out
