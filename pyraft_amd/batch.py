"""Batches of independent slices with different N, R and A (libraft ``raft_ss_batch_*``).

The reference runs one slice per call (``raft_radon_slantstack_gpu``,
raft/raft_radon_gpu.cu:16-35; ``raft_em_gpu``, raft/raft_em_gpu.cu:20-88) and loops
over OSEM subsets in Python (pyraft/iterative/em.py:265-320).  A ``Batch`` runs all
slices of a batch in one launch per kernel (persistent, heaviest items first, for the
Radon and backprojection), and with ``graph=True`` the whole call is captured once
into a HIP graph and replayed by one graph launch.

Data live in packed float32 device buffers: slice i's image is the n_i x n_i block
at ``img_off[i]`` and its sinogram the R_i x A_i block (or A_i x R_i, angle-major)
at ``sino_off[i]``; ``pack_images`` / ``unpack_images`` and the sinogram
equivalents convert from and to lists of tensors.
"""
import ctypes

import torch

from .device import DEFAULT_IMG_CORNERS, DEFAULT_SINO_CORNERS, _require_gpu, current_stream_handle
from .raftypes import BATCH_GRAPH, check, libraft


class Batch:
    def __init__(self, sizes, nrays=None, nangles=None, img_corners=None, sino_corners=None, device=None):
        """``sizes``: image sides n_i; ``nrays`` (default n_i) and ``nangles`` per slice.
        Corners: None (defaults) or one (tl_x, tl_y, br_x, br_y) per slice."""
        _require_gpu()
        self.n = [int(v) for v in sizes]
        ns = len(self.n)
        if ns == 0:
            raise ValueError("a batch needs at least one slice")
        self.nrays = [int(v) for v in (nrays if nrays is not None else self.n)]
        if nangles is None:
            raise ValueError("nangles is required")
        self.nangles = [int(v) for v in nangles]
        if not (len(self.nrays) == len(self.nangles) == ns):
            raise ValueError("sizes, nrays and nangles must have one entry per slice")
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else
                                   torch.device(device).index or 0)
        ia = (ctypes.c_int * ns)
        ic = sc = None
        if img_corners is not None:
            ic = (ctypes.c_double * (4 * ns))(*[float(v) for c in img_corners for v in c])
        if sino_corners is not None:
            sc = (ctypes.c_double * (4 * ns))(*[float(v) for c in sino_corners for v in c])
        self._handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(libraft.raft_ss_batch_create(ctypes.byref(self._handle), ns, ia(*self.n), ia(*self.nrays),
                                               ia(*self.nangles), ic, sc, current_stream_handle()),
                  "raft_ss_batch_create")
        io = (ctypes.c_longlong * ns)()
        so = (ctypes.c_longlong * ns)()
        tot = (ctypes.c_longlong * 2)()
        check(libraft.raft_ss_batch_layout(self._handle, io, so, tot), "raft_ss_batch_layout")
        self.img_off = list(io)
        self.sino_off = list(so)
        self.img_total, self.sino_total = int(tot[0]), int(tot[1])

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            try:
                libraft.raft_ss_batch_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._handle = None

    def __len__(self):
        return len(self.n)

    def _stream(self):
        return current_stream_handle(self.device)

    # -- packing
    def empty_images(self):
        return torch.zeros(self.img_total, dtype=torch.float32, device=self.device)

    def empty_sinos(self):
        return torch.zeros(self.sino_total, dtype=torch.float32, device=self.device)

    def image(self, packed, i):
        """View of slice i's image in a packed buffer."""
        n = self.n[i]
        return packed[self.img_off[i]:self.img_off[i] + n * n].view(n, n)

    def sino(self, packed, i, angle_major=False):
        """View of slice i's sinogram in a packed buffer."""
        R, A = self.nrays[i], self.nangles[i]
        v = packed[self.sino_off[i]:self.sino_off[i] + R * A]
        return v.view(A, R) if angle_major else v.view(R, A)

    def pack_images(self, images):
        out = self.empty_images()
        for i, img in enumerate(images):
            self.image(out, i).copy_(img)
        return out

    def pack_sinos(self, sinos, angle_major=False):
        out = self.empty_sinos()
        for i, s in enumerate(sinos):
            self.sino(out, i, angle_major).copy_(s)
        return out

    def unpack_images(self, packed):
        return [self.image(packed, i) for i in range(len(self))]

    def unpack_sinos(self, packed, angle_major=False):
        return [self.sino(packed, i, angle_major) for i in range(len(self))]

    # -- configuration
    def info(self, i):
        out = (ctypes.c_int * 6)()
        check(libraft.raft_ss_batch_info(self._handle, int(i), out, 6), "raft_ss_batch_info")
        keys = ("n", "nrays", "nangles", "irregular_rays", "bp_variant", "bp_angles")
        return dict(zip(keys, list(out)))

    def set_bp_variant(self, variant, slice_index=-1):
        check(libraft.raft_ss_batch_set_bp_variant(self._handle, int(slice_index), int(variant)),
              "raft_ss_batch_set_bp_variant")

    def tune(self, reps=2):
        """Autotune the backprojection variant per image size; returns the variant per slice."""
        with torch.cuda.device(self.device):
            check(libraft.raft_ss_batch_tune(self._handle, int(reps), 0, self._stream()), "raft_ss_batch_tune")
        return [self.info(i)["bp_variant"] for i in range(len(self))]

    # -- operators (packed float32 device buffers)
    def _packed(self, t, total, name):
        if not isinstance(t, torch.Tensor) or t.dtype != torch.float32 or t.device.type != "cuda":
            raise TypeError("%s must be a float32 device tensor" % name)
        if not t.is_contiguous() or t.numel() < total:
            raise ValueError("%s must be contiguous with at least %d elements" % (name, total))
        return t

    def radon(self, images, angle_major=False, out=None, graph=False):
        x = self._packed(images, self.img_total, "images")
        out = self.empty_sinos() if out is None else self._packed(out, self.sino_total, "out")
        with torch.cuda.device(self.device):
            check(libraft.raft_ss_batch_radon(self._handle, x.data_ptr(), out.data_ptr(), int(angle_major),
                                              BATCH_GRAPH if graph else 0, self._stream()), "raft_ss_batch_radon")
        return out

    def backprojection(self, sinos, angle_major=False, out=None, graph=False):
        s = self._packed(sinos, self.sino_total, "sinos")
        out = self.empty_images() if out is None else self._packed(out, self.img_total, "out")
        with torch.cuda.device(self.device):
            check(libraft.raft_ss_batch_backprojection(self._handle, s.data_ptr(), out.data_ptr(), int(angle_major),
                                                       BATCH_GRAPH if graph else 0, self._stream()),
                  "raft_ss_batch_backprojection")
        return out

    def em(self, sinos, niter, nsubsets=1, x=None, return_ll=False, graph=False):
        """OSEM of every slice from x (default ones), in place on ``x`` when given.
        Returns x (and the log-likelihood, a niter x nslices list of lists)."""
        s = self._packed(sinos, self.sino_total, "sinos")
        if x is None:
            x = torch.ones(self.img_total, dtype=torch.float32, device=self.device)
        else:
            x = self._packed(x, self.img_total, "x")
        ns = len(self)
        ll = (ctypes.c_double * max(1, int(niter) * ns))()
        with torch.cuda.device(self.device):
            check(libraft.raft_ss_batch_em(self._handle, s.data_ptr(), x.data_ptr(), int(niter), int(nsubsets),
                                           ll if return_ll else None, BATCH_GRAPH if graph else 0, self._stream()),
                  "raft_ss_batch_em")
        if return_ll:
            return x, [[ll[it * ns + i] for i in range(ns)] for it in range(int(niter))]
        return x
