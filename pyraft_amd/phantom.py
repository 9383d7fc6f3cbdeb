"""Ellipse phantoms (inputs for the projector path).

``shepplogan`` mirrors pyraft/phantom/phantom.py:12-52: the same 10-ellipse table
(columns value, a, b, x0, y0, rotation) rasterised by ``raft_image_phantom_fromdesc``
(raft/raft_image_phantom_fromdesc.c:4-59), row 0 at the image's top-left corner.

``random_ellipses`` draws the seeded random-ellipse descriptions used for the
large workloads: centres uniform in the unit disk (r = sqrt(U), phi = U(0, 2 pi)),
semi-axes U(0.02, 0.4), densities U(0, 1), rotations U(0, pi), additive.
"""
import math

import numpy

from .raftypes import check_legacy, image, libraft, make_RAFT_IMAGE, make_RAFT_MATRIX

__all__ = ["ellipses_phantom", "shepplogan", "random_ellipses"]


def _shepp_logan_desc():
    """The Shepp-Logan head description (pyraft/phantom/phantom.py:15-24)."""
    p10 = math.pi / 10.0
    return numpy.array([[1.0, 0.69, 0.92, 0.0, 0.0, 0.0],
                        [-0.8, 0.6624, 0.874, 0.0, -0.0184, 0.0],
                        [-0.2, 0.1100, 0.31, 0.22, 0.0, -p10],
                        [-0.2, 0.1600, 0.41, -0.22, 0.0, p10],
                        [0.1, 0.2100, 0.25, 0.0, 0.35, 0.0],
                        [0.1, 0.0460, 0.046, 0.0, 0.1, 0.0],
                        [0.1, 0.0460, 0.046, 0.0, -0.1, 0.0],
                        [0.1, 0.0460, 0.023, -0.08, -0.605, 0.0],
                        [0.1, 0.0230, 0.023, 0.0, -0.606, 0.0],
                        [0.1, 0.0230, 0.046, 0.06, -0.605, 0.0]])


def ellipses_phantom(desc, shape=(256, 256), **kwargs):
    """Rasterise an ellipse description into a float64 ``image`` (additive)."""
    desc = numpy.ascontiguousarray(desc, dtype=numpy.float64)
    if desc.ndim != 2 or desc.shape[1] < 6:
        raise ValueError("ellipse description must be K x 6 (value, a, b, x0, y0, rotation)")
    img = image(shape, **kwargs)
    libraft.raft_image_phantom_fromdesc(make_RAFT_IMAGE(img, img.top_left, img.bottom_right),
                                        make_RAFT_MATRIX(desc))
    check_legacy("raft_image_phantom_fromdesc")
    return img


def shepplogan(shape=(256, 256), **kwargs):
    """Shepp-Logan head phantom (pyraft/phantom/phantom.py:33-52)."""
    return ellipses_phantom(_shepp_logan_desc(), shape, **kwargs)


def random_ellipses(count=32, seed=0):
    """Seeded K x 6 random-ellipse description (numpy PCG64)."""
    rng = numpy.random.default_rng(seed)
    r = numpy.sqrt(rng.random(count))
    phi = rng.uniform(0.0, 2.0 * math.pi, count)
    a = rng.uniform(0.02, 0.4, count)
    b = rng.uniform(0.02, 0.4, count)
    val = rng.random(count)
    rot = rng.uniform(0.0, math.pi, count)
    return numpy.stack([val, a, b, r * numpy.cos(phi), r * numpy.sin(phi), rot], axis=1)
