"""FBP ramp filters (pyraft/filters/filters.py:13-111).

``lowpassino_fbp(sino[, t])`` applies, to every angle column, the filter
Re(ifft(fft(s) * fftshift(h))) with h = |w| cos(pi w dt), w = linspace(-wc, wc, R),
wc = 1/(2 dt).  Because s is real this is exactly a circular convolution along t
with the real kernel g = Re(ifft(fftshift(h))); libraft computes g in FP64 and runs
the convolution as a batched row filter on the GPU.  ``lowpassino_reg_fbp(sino, L)``
is the same with h = |w| / (1 + L |w|^2).
"""
import numpy

from .raftypes import image

__all__ = ["lowpassino_fbp", "lowpassino_reg_fbp"]


def lowpassino_fbp(*args):
    if len(args) == 0:
        raise TypeError("lowpassino_fbp needs a sinogram")
    sino = args[0]
    tl = getattr(sino, 'top_left', None) or (0.0, 1.0)
    br = getattr(sino, 'bottom_right', None) or (numpy.pi, -1.0)
    R = int(sino.shape[0])
    if len(args) > 1:
        t = numpy.asarray(args[1], dtype=numpy.float64)
        tmin, tmax = float(t[0]), float(t[R - 1])
    else:
        tmin, tmax = float(br[1]), float(tl[1])
    return _filter(sino, tl, br, tmin, tmax, 0, 0.0)


def lowpassino_reg_fbp(*args):
    """lowpassino_reg_fbp(sino, L) (pyraft/filters/filters.py:73-111): the same circular
    filter with the Tikhonov-damped ramp h = |w| / (1 + L |w|^2).  L is required (pyraft
    reads it only from a second argument)."""
    if len(args) == 0:
        raise TypeError("lowpassino_reg_fbp needs a sinogram")
    if len(args) < 2:
        raise TypeError("lowpassino_reg_fbp needs the regularisation parameter L")
    sino = args[0]
    tl = getattr(sino, 'top_left', None) or (0.0, 1.0)
    br = getattr(sino, 'bottom_right', None) or (numpy.pi, -1.0)
    return _filter(sino, tl, br, float(br[1]), float(tl[1]), 1, float(args[1]))


def _filter(sino, tl, br, tmin, tmax, kind, param):
    import torch
    from .device import filter_rows, ramp_kernel
    R = int(sino.shape[0])
    g = torch.from_numpy(ramp_kernel(R, tmin, tmax, kind, param)).cuda()
    rows = torch.from_numpy(numpy.ascontiguousarray(numpy.asarray(sino, dtype=numpy.float32).T)).cuda()
    out = filter_rows(rows, g).cpu().numpy().T.astype(numpy.float64)
    W = image(numpy.ascontiguousarray(out))
    W.top_left = tl
    W.bottom_right = br
    return W
