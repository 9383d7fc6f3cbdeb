"""ctypes boundary to ``libraft`` and the ``image`` array type.

Mirrors pyraft/raftypes.py: the shared library is loaded from
``<package>/lib/libraft<EXT_SUFFIX>`` (pyraft/raftypes.py:29-38), the by-value
structs ``RAFT_MATRIX`` / ``RAFT_IMAGE`` / ``RAFT_VECTOR`` mirror
raft/raft_matrix.h:11-19, raft/raft_image.h:11-19 and raft/raft_vector.h:10-16
(pyraft/raftypes.py:85-107), and ``image`` is an ndarray that carries its field of
view corners (pyraft/raftypes.py:254-406).

Differences from the reference, all deliberate:
* a missing library raises ImportError (the reference fails the same way; there is
  no CPU fallback here);
* ``make_RAFT_MATRIX`` rejects non-float64 arrays instead of reinterpreting their
  bytes as doubles (pyraft/raftypes.py:415 casts any dtype to ``double*``).
"""
import ctypes
import multiprocessing
import os
import sysconfig

import numpy

__all__ = ["image_read", "image_write", "nthreads", "libraft", "LIBRARY_PATH", "HIP_RUNTIME", "RAFT_MATRIX", "RAFT_IMAGE", "RAFT_VECTOR", "RaftError",
           "check", "check_legacy", "image", "make_RAFT_MATRIX", "make_RAFT_IMAGE", "make_RAFT_VECTOR",
           "EXPORTED_SYMBOLS"]

nthreads = multiprocessing.cpu_count()

_LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
_EXT = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
# PYRAFT_AMD_LIBRARY selects another build of the library (e.g. a locally tuned one)
LIBRARY_PATH = os.environ.get("PYRAFT_AMD_LIBRARY") or os.path.join(_LIBDIR, "libraft" + _EXT)

if not os.path.exists(LIBRARY_PATH):
    raise ImportError(
        "pyraft_amd: native library %s is not built; run __graft_entry__.build() "
        "(or python pyraft_amd/build.py)" % LIBRARY_PATH)


def _preload_hip_runtime():
    """One HIP runtime per process.

    ROCm PyTorch wheels carry their own libamdhip64 / libhsa-runtime64 and link them
    by the bare file name, which does not match the SONAME of the system copy that
    libraft links.  If libraft were loaded first, importing torch afterwards would
    bring up a second HIP + HSA runtime and one of them would find no device.  Loading
    torch's copy first makes libraft's libamdhip64.so.7 dependency resolve to it (same
    SONAME), whatever the import order.  PYRAFT_AMD_HIP_RUNTIME selects another file."""
    cands = []
    if os.environ.get("PYRAFT_AMD_HIP_RUNTIME"):
        cands.append(os.environ["PYRAFT_AMD_HIP_RUNTIME"])
    else:
        try:
            import importlib.util
            spec = importlib.util.find_spec("torch")
            for d in (spec.submodule_search_locations or []) if spec else []:
                cands.append(os.path.join(d, "lib", "libamdhip64.so"))
        except (ImportError, ValueError):
            pass
    for c in cands:
        if os.path.exists(c):
            ctypes.CDLL(c, mode=ctypes.RTLD_GLOBAL)
            return c
    return None


HIP_RUNTIME = _preload_hip_runtime()
libraft = ctypes.CDLL(LIBRARY_PATH)

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_float_p = ctypes.POINTER(ctypes.c_float)


class RAFT_MATRIX(ctypes.Structure):
    """raft_matrix: strided FP64 view, element (i,j) at p_data[i*column_stride + j*line_stride]."""
    _fields_ = [("p_data", _c_double_p),
                ("lines", ctypes.c_int),
                ("line_stride", ctypes.c_int),
                ("columns", ctypes.c_int),
                ("column_stride", ctypes.c_int)]


class RAFT_IMAGE(ctypes.Structure):
    """raft_image: raft_matrix plus its field-of-view corners."""
    _fields_ = [("data", RAFT_MATRIX),
                ("tl_x", ctypes.c_double),
                ("tl_y", ctypes.c_double),
                ("br_x", ctypes.c_double),
                ("br_y", ctypes.c_double)]


class RAFT_VECTOR(ctypes.Structure):
    _fields_ = [("p_data", _c_double_p),
                ("size", ctypes.c_int),
                ("stride", ctypes.c_int)]


def _decl(name, argtypes, restype):
    fn = getattr(libraft, name)
    fn.argtypes = argtypes
    fn.restype = restype
    return fn


_vp = ctypes.c_void_p
_i = ctypes.c_int
_d4 = ctypes.POINTER(ctypes.c_double)

# drop-in symbols (reference signatures, pyraft/raftypes.py:138-215)
_decl("raft_image_phantom_fromdesc", [RAFT_IMAGE, RAFT_MATRIX], None)
_decl("raft_radon_fromdesc", [RAFT_IMAGE, RAFT_MATRIX], None)
_decl("raft_radon_slantstack", [RAFT_IMAGE, RAFT_IMAGE, _i], None)
_decl("raft_backprojection_slantstack", [RAFT_IMAGE, RAFT_IMAGE, _i], None)
_decl("raft_radon_slantstack_gpu", [_c_float_p, _c_float_p, _i, _i, _i], None)
_decl("raft_backprojection_slantstack_gpu", [_c_float_p, _c_float_p, _i, _i, _i], None)
_decl("raft_em_gpu", [_c_float_p, _c_float_p, _i, _i, _i, _i], None)
_decl("raft_radon_transpose_slantstack", [RAFT_IMAGE, RAFT_IMAGE, _i], None)
_decl("raft_radon_bresenham", [RAFT_IMAGE, RAFT_IMAGE, _i], None)
_decl("raft_radon_transpose_bresenham", [RAFT_IMAGE, RAFT_IMAGE, _i], None)
_decl("raft_backprojection_bresenham", [RAFT_IMAGE, RAFT_IMAGE, _i], None)
_decl("raft_tr_em_gpu", [_c_float_p, _c_float_p, _c_float_p, _i, _i, _i, _i], None)
_decl("raft_matrix_read", [_vp], RAFT_MATRIX)
_decl("raft_matrix_write", [RAFT_MATRIX, _vp], _i)
_decl("raft_matrix_destroy", [ctypes.POINTER(RAFT_MATRIX)], None)
_decl("raft_image_read", [_vp], RAFT_IMAGE)
_decl("raft_image_write", [RAFT_IMAGE, _vp], _i)
_decl("raft_image_destroy", [ctypes.POINTER(RAFT_IMAGE)], None)
_decl("raft_filter1d_ramp", [ctypes.c_double, RAFT_IMAGE, _i], None)
_decl("raft_fbp", [ctypes.c_double, RAFT_IMAGE, RAFT_IMAGE, _i, _i], None)
_decl("raft_radon_slantstack_withmesh", [RAFT_IMAGE, RAFT_IMAGE, RAFT_VECTOR, RAFT_VECTOR, _i], None)
_decl("raft_radon_slantstack_view", [RAFT_VECTOR, RAFT_IMAGE, RAFT_VECTOR, ctypes.c_double], None)
_decl("raft_ss_legacy_status", [], _i)

# new entry points
_decl("raft_ss_last_error", [], ctypes.c_char_p)
_decl("raft_ss_version", [], _i)
_decl("raft_ss_plan_create", [ctypes.POINTER(_vp), _i, _i, _i, _d4, _d4, _vp], _i)
_decl("raft_ss_plan_destroy", [_vp], _i)
_decl("raft_ss_plan_info", [_vp, ctypes.POINTER(_i), _i], _i)
_decl("raft_ss_plan_bounds", [_vp, ctypes.POINTER(_i)], _i)
_decl("raft_ss_plan_wave_bounds", [_vp, ctypes.POINTER(_i)], _i)
_decl("raft_ss_plan_host_bounds", [_i, _i, _i, _d4, _d4, ctypes.POINTER(_i),
                                   ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(_i)], _i)
_decl("raft_ss_plan_host_info", [_i, _i, _i, _d4, _d4, ctypes.POINTER(_i)], _i)
_decl("raft_dda_plan_create", [ctypes.POINTER(_vp), _i, _i, _i, _i, _d4, _d4, _vp], _i)
_decl("raft_dda_plan_destroy", [_vp], _i)
_decl("raft_dda_plan_info", [_vp, ctypes.POINTER(_i), _i], _i)
_decl("raft_dda_radon", [_vp, _vp, _vp, _i, _vp], _i)
_decl("raft_dda_radon_transpose", [_vp, _vp, _vp, _i, _vp], _i)
_decl("raft_dda_backprojection", [_vp, _vp, _vp, _i, _vp], _i)
_decl("raft_dda_weights", [_vp, _vp, _vp], _i)
_decl("raft_dda_host_ray", [_i, _i, _i, _i, _d4, _d4, _i, _i, ctypes.POINTER(_i), _d4, _i], _i)
_decl("raft_ss_plan_create_mesh", [ctypes.POINTER(_vp), _i, _d4, _i, _d4, _i, _d4, _vp], _i)
_decl("raft_ss_plan_host_bounds_mesh", [_i, _d4, _i, _d4, _i, _d4, ctypes.POINTER(_i),
                                        ctypes.POINTER(ctypes.c_longlong)], _i)
_decl("raft_ss_radon", [_vp, _vp, _vp, _i, _vp], _i)
_decl("raft_ss_radon_range", [_vp, _vp, _vp, _i, _i, _i, _i, _vp], _i)
# allocation-free Radon (include/raft_noalloc.h): workspace in a static device array
_decl("raft_ss_radon_noalloc", [_vp, _vp, _vp, _i, _i, _i, _i, _vp], _i)
_decl("raft_ss_backprojection", [_vp, _vp, _vp, _i, _vp], _i)
_decl("raft_ss_backprojection_range", [_vp, _vp, _vp, _i, _i, _i, _i, _vp], _i)
_decl("raft_ss_radon_transpose", [_vp, _vp, _vp, _i, _vp], _i)
_decl("raft_ss_radon_transpose_range", [_vp, _vp, _vp, _i, _i, _i, _i, _vp], _i)
_decl("raft_ss_fbp_filter", [_vp, _vp, _vp, _i, _i, _vp], _i)
_decl("raft_ss_fbp", [_vp, _vp, _vp, _i, _vp], _i)
_decl("raft_ss_em", [_vp, _vp, _vp, _i, _i, _i, _d4, _vp], _i)
_decl("raft_ss_tr_em", [_vp, _vp, _vp, _vp, _i, _i, _vp], _i)
_decl("raft_ss_phantom", [_d4, _i, _i, _i, _d4, _vp, _vp], _i)
_decl("raft_ss_ramp_kernel", [_i, ctypes.c_double, ctypes.c_double, _c_float_p], _i)
_decl("raft_ss_filter_rows", [_vp, _vp, _vp, _i, _i, _vp], _i)
_decl("raft_ss_filter_kernel", [_i, ctypes.c_double, ctypes.c_double, _i, ctypes.c_double, _c_float_p], _i)
_decl("raft_ss_filter1d_ramp_taps", [_i, ctypes.c_double, ctypes.c_double, _c_float_p], _i)
_decl("raft_ss_filter_rows_toeplitz", [_vp, _vp, _vp, _i, _i, _vp], _i)
_decl("raft_ss_plan_host_bp_staging", [_i, _i, _i, _d4, _d4, ctypes.POINTER(_i), ctypes.POINTER(_i)], _i)
_decl("raft_ss_plan_host_rt_staging", [_i, _i, _i, _d4, _d4, ctypes.POINTER(_i), ctypes.POINTER(_i), _vp], _i)
_decl("raft_ss_bp_variant_shape", [_i, ctypes.POINTER(_i), ctypes.POINTER(_i)], _i)
_decl("raft_ss_plan_set_bp_variant", [_vp, _i], _i)
_decl("raft_ss_plan_set_radon_slab_rows", [_vp, _i], _i)
_decl("raft_ss_plan_tune_transpose", [_vp, _i, _i, _i, _d4, _vp], _i)
_decl("raft_ss_plan_tune", [_vp, _i, _i, _d4, _vp], _i)
_decl("raft_ss_batch_create", [ctypes.POINTER(_vp), _i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                               _d4, _d4, _vp], _i)
_decl("raft_ss_batch_destroy", [_vp], _i)
_decl("raft_ss_batch_layout", [_vp, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                               ctypes.POINTER(ctypes.c_longlong)], _i)
_decl("raft_ss_batch_info", [_vp, _i, ctypes.POINTER(_i), _i], _i)
_decl("raft_ss_batch_set_bp_variant", [_vp, _i, _i], _i)
_decl("raft_ss_batch_tune", [_vp, _i, _i, _vp], _i)
_decl("raft_ss_batch_radon", [_vp, _vp, _vp, _i, _i, _vp], _i)
_decl("raft_ss_batch_backprojection", [_vp, _vp, _vp, _i, _i, _vp], _i)
_decl("raft_ss_batch_em", [_vp, _vp, _vp, _i, _i, _d4, _i, _vp], _i)

BP_NVARIANTS = 6          # RAFT_SS_BP_NVARIANTS
TUNE_FORCE = 1            # RAFT_SS_TUNE_FORCE
BATCH_GRAPH = 1           # RAFT_SS_BATCH_GRAPH


class RaftError(RuntimeError):
    """A raft_ss_* entry point returned a non-zero status."""


def check_legacy(what):
    """Raise if the last drop-in (void) libraft call on this thread failed."""
    check(libraft.raft_ss_legacy_status(), what)


def check(rc, what=""):
    if rc != 0:
        msg = libraft.raft_ss_last_error()
        msg = msg.decode() if msg else "unknown error"
        raise RaftError("%s: %s (status %d)" % (what or "libraft", msg, rc))


# ------------------------------------------------------------------ image

class image(numpy.ndarray):
    """ndarray carrying the corners of its field of view.

    Same constructor contract as pyraft.image (pyraft/raftypes.py:258-358):
    ``image(shape | ndarray | image, top_left=, bottom_right=, extent=, x_extent=, y_extent=)``;
    default corners (-1, 1) / (1, -1); a bare shape gives a zero float64 array."""

    def __new__(subtype, shape, top_left=None, bottom_right=None, extent=None,
                x_extent=None, y_extent=None, **kwargs):
        extent_given = extent or x_extent or y_extent
        corner_given = top_left or bottom_right
        if extent_given and corner_given:
            raise TypeError('Mutually exclusive arguments given.')
        if extent_given:
            if not extent:
                x_extent = x_extent or (-1.0, 1.0)
                y_extent = y_extent or (-1.0, 1.0)
            else:
                x_extent = (extent[0], extent[1])
                y_extent = (extent[2], extent[3])
            top_left = (x_extent[0], y_extent[1])
            bottom_right = (x_extent[1], y_extent[0])

        if isinstance(shape, image):
            if not extent_given:
                top_left = top_left or shape.top_left
                bottom_right = bottom_right or shape.bottom_right
            if kwargs:
                raise TypeError('Unhandled arguments!')
            obj = numpy.ndarray.__new__(subtype, shape.shape, dtype=shape.dtype)
            obj[...] = shape[...]
        else:
            if not extent_given:
                top_left = top_left or (-1.0, 1.0)
                bottom_right = bottom_right or (1.0, -1.0)
            if isinstance(shape, numpy.ndarray):
                if kwargs:
                    raise TypeError('Unhandled arguments!')
                obj = numpy.ndarray.__new__(subtype, shape.shape, dtype=shape.dtype)
                obj[...] = shape[...]
            else:
                kwargs.setdefault('dtype', numpy.float64)
                obj = numpy.ndarray.__new__(subtype, shape, **kwargs)
                obj[...] = 0.0

        if len(obj.shape) != len(top_left) or len(top_left) != len(bottom_right):
            raise TypeError('Dimensions must match!')
        obj.top_left = tuple(top_left)
        obj.bottom_right = tuple(bottom_right)
        try:
            obj.sampling_distances = ((bottom_right[0] - top_left[0]) / (obj.shape[1] - 1.0),
                                      (bottom_right[1] - top_left[1]) / (obj.shape[0] - 1.0))
        except ZeroDivisionError:
            obj.sampling_distances = (0.0, 0.0)
        return obj

    def __array_finalize__(self, obj):
        if obj is None:
            return
        self.top_left = getattr(obj, 'top_left', None)
        self.bottom_right = getattr(obj, 'bottom_right', None)
        self.sampling_distances = getattr(obj, 'sampling_distances', None)

    def __reduce__(self):
        state = list(numpy.ndarray.__reduce__(self))
        state[2] = (state[2], (self.top_left, self.bottom_right, self.sampling_distances))
        return tuple(state)

    def __setstate__(self, state):
        numpy.ndarray.__setstate__(self, state[0])
        self.top_left, self.bottom_right, self.sampling_distances = state[1]

    def sample_coordinates(self, idx):
        return (self.top_left[0] + idx[1] * self.sampling_distances[0],
                self.top_left[1] + idx[0] * self.sampling_distances[1])

    def get_y_coordinate(self, idx):
        return self.top_left[1] + idx * self.sampling_distances[1]

    def get_x_coordinate(self, idx):
        return self.top_left[0] + idx * self.sampling_distances[0]

    def extent(self):
        return (self.top_left[0], self.bottom_right[0], self.bottom_right[1], self.top_left[1])


def make_RAFT_MATRIX(array):
    """raft_matrix view of a 2-D float64 array (pyraft/raftypes.py:413-420)."""
    if not isinstance(array, numpy.ndarray) or array.ndim != 2:
        raise TypeError('make_RAFT_MATRIX needs a 2-D numpy array')
    if array.dtype != numpy.float64:
        raise TypeError('make_RAFT_MATRIX needs float64 data (got %s)' % array.dtype)
    if array.strides[0] % 8 or array.strides[1] % 8:
        raise TypeError('make_RAFT_MATRIX needs element-aligned strides')
    return RAFT_MATRIX(ctypes.cast(array.ctypes.data, _c_double_p),
                       ctypes.c_int(int(array.shape[0])),
                       ctypes.c_int(int(array.strides[1] // 8)),
                       ctypes.c_int(int(array.shape[1])),
                       ctypes.c_int(int(array.strides[0] // 8)))


def make_RAFT_IMAGE(array, top_left=(-1.0, 1.0), bottom_right=(1.0, -1.0)):
    """raft_image from an ndarray / pyraft image / RAFT_MATRIX (pyraft/raftypes.py:427-449)."""
    if isinstance(array, RAFT_MATRIX):
        mat = array
    elif isinstance(array, numpy.ndarray):
        mat = make_RAFT_MATRIX(array)
    else:
        raise TypeError('make_RAFT_IMAGE needs an ndarray or RAFT_MATRIX')
    return RAFT_IMAGE(mat, ctypes.c_double(top_left[0]), ctypes.c_double(top_left[1]),
                      ctypes.c_double(bottom_right[0]), ctypes.c_double(bottom_right[1]))


def make_RAFT_VECTOR(array):
    if array.dtype != numpy.float64:
        raise TypeError('make_RAFT_VECTOR needs float64 data')
    return RAFT_VECTOR(ctypes.cast(array.ctypes.data, _c_double_p),
                       ctypes.c_int(int(array.shape[0])),
                       ctypes.c_int(int(array.strides[0] // 8)))


def as_float64_image(arr, top_left=None, bottom_right=None):
    """float64 ``image`` with the given (or the array's own, or default) corners."""
    tl = top_left or getattr(arr, 'top_left', None) or (-1.0, 1.0)
    br = bottom_right or getattr(arr, 'bottom_right', None) or (1.0, -1.0)
    out = numpy.asarray(arr, dtype=numpy.float64)
    out = image(out, top_left=tuple(tl), bottom_right=tuple(br))
    return out


def corners(arr, default_tl=(-1.0, 1.0), default_br=(1.0, -1.0)):
    tl = getattr(arr, 'top_left', None) or default_tl
    br = getattr(arr, 'bottom_right', None) or default_br
    return (float(tl[0]), float(tl[1]), float(br[0]), float(br[1]))


EXPORTED_SYMBOLS = (
    "raft_radon_slantstack", "raft_backprojection_slantstack", "raft_radon_slantstack_gpu",
    "raft_backprojection_slantstack_gpu", "raft_em_gpu", "raft_image_phantom_fromdesc",
    "raft_radon_fromdesc", "raft_radon_transpose_slantstack", "raft_ss_legacy_status", "raft_ss_last_error", "raft_ss_version", "raft_ss_plan_create",
    "raft_ss_plan_destroy", "raft_ss_plan_info", "raft_ss_plan_bounds", "raft_ss_plan_wave_bounds", "raft_ss_plan_host_bounds", "raft_ss_plan_host_info",
    "raft_ss_radon", "raft_ss_radon_range", "raft_ss_radon_noalloc", "raft_ss_backprojection", "raft_ss_backprojection_range",
    "raft_ss_radon_transpose", "raft_ss_radon_transpose_range", "raft_ss_fbp_filter", "raft_ss_fbp",
    "raft_ss_em", "raft_ss_phantom", "raft_ss_ramp_kernel", "raft_ss_filter_rows",
    "raft_ss_plan_host_bp_staging", "raft_ss_bp_variant_shape", "raft_ss_plan_set_bp_variant", "raft_ss_plan_tune",
    "raft_ss_batch_create", "raft_ss_batch_destroy", "raft_ss_batch_layout", "raft_ss_batch_info",
    "raft_ss_batch_set_bp_variant", "raft_ss_batch_tune", "raft_ss_batch_radon", "raft_ss_batch_backprojection",
    "raft_ss_batch_em", "raft_radon_slantstack_withmesh", "raft_radon_slantstack_view",
    "raft_ss_plan_create_mesh", "raft_ss_plan_host_bounds_mesh",
    "raft_radon_bresenham", "raft_radon_transpose_bresenham", "raft_backprojection_bresenham",
    "raft_dda_plan_create", "raft_dda_plan_destroy", "raft_dda_plan_info", "raft_dda_radon",
    "raft_dda_radon_transpose", "raft_dda_backprojection", "raft_dda_weights", "raft_dda_host_ray",
    "raft_filter1d_ramp", "raft_fbp", "raft_ss_filter_kernel", "raft_ss_filter1d_ramp_taps",
    "raft_ss_filter_rows_toeplitz", "raft_tr_em_gpu", "raft_ss_tr_em", "raft_matrix_read", "raft_matrix_write",
    "raft_matrix_destroy", "raft_image_read", "raft_image_write", "raft_image_destroy",
)
RAFT_BRESENHAM, RAFT_SLANTSTACK = 0, 1          # raft_radon_method (raft/raft_enums.h:4-9)
FILTER_COS_RAMP, FILTER_TIKHONOV_RAMP = 0, 1    # raft_ss_filter_kernel kinds


# ---------------------------------------------------------------- binary image files

_libc = ctypes.CDLL(None)
_libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
_libc.fopen.restype = ctypes.c_void_p
_libc.fclose.argtypes = [ctypes.c_void_p]
_libc.fclose.restype = ctypes.c_int


def image_write(img, path):
    """Write a pyraft image in the reference's binary layout (raft_image_write,
    raft/raft_image.c:135-151): int32 lines, int32 columns, FP64 values, 4 FP64 corners."""
    a = numpy.asarray(img, dtype=numpy.float64)
    tl = tuple(getattr(img, 'top_left', (-1.0, 1.0)))
    br = tuple(getattr(img, 'bottom_right', (1.0, -1.0)))
    fp = _libc.fopen(os.fsencode(path), b"wb")
    if not fp:
        raise OSError("cannot open %s for writing" % path)
    try:
        rc = libraft.raft_image_write(make_RAFT_IMAGE(a, tl, br), fp)
    finally:
        _libc.fclose(fp)
    if rc:
        raise OSError("raft_image_write failed (%d) on %s" % (rc, path))


def image_read(path):
    """Read a binary raft image file (raft_image_read, raft/raft_image.c:5-38) into a
    pyraft image with its corners."""
    fp = _libc.fopen(os.fsencode(path), b"rb")
    if not fp:
        raise OSError("cannot open %s" % path)
    try:
        im = libraft.raft_image_read(fp)
    finally:
        _libc.fclose(fp)
    if not im.data.p_data:
        raise OSError("%s is not a complete raft image file" % path)
    try:
        n = im.data.lines * im.data.columns
        vals = numpy.ctypeslib.as_array(im.data.p_data, shape=(n,)).reshape(im.data.lines, im.data.columns).copy()
        return image(vals, top_left=(im.tl_x, im.tl_y), bottom_right=(im.br_x, im.br_y))
    finally:
        libraft.raft_image_destroy(ctypes.byref(im))
