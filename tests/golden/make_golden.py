"""Write tests/golden/slantstack_golden.npz from the reference's own C.

Runs oracle/_ref/libraft_ref.so, which oracle/Makefile compiles from the reference's
sources (raft/raft_radon_slantstack.c, raft/raft_backprojection_slantstack.c,
raft/raft_radon_slantstack_withmesh.c, raft/raft_radon_slantstack_view.c,
raft/raft_image_phantom_fromdesc.c, raft/raft_radon_fromdesc.c, raft/raft_ddaiterator.c and
their helpers), on small seeded inputs.  The fixture holds inputs and outputs only.

    make -C oracle ref && python tests/golden/make_golden.py

It also writes image_written_by_reference.raft with the reference's raft_image_write.
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

DEF_IMG = (-1.0, 1.0, 1.0, -1.0)
DEF_SINO = (0.0, 1.0, math.pi, -1.0)

# (name, n, R, A, image corners, sinogram corners)
CASES = [
    ("n8_r8_a6", 8, 8, 6, DEF_IMG, DEF_SINO),
    ("n17_r20_a45", 17, 20, 45, DEF_IMG, DEF_SINO),
    ("n64_r64_a45", 64, 64, 45, DEF_IMG, DEF_SINO),
    ("n64_r67_a180", 64, 67, 180, DEF_IMG, DEF_SINO),
    ("n33_r40_a7_fov", 33, 40, 7, (-0.7, 1.3, 1.2, -0.9), (-math.pi / 2, 1.1, math.pi / 2, -0.8)),
    ("n48_r48_a4", 48, 48, 4, DEF_IMG, DEF_SINO),
    ("n65_r64_a90_shift", 65, 64, 90, DEF_IMG, (0.3, 1.0, 2.5, -1.0)),
]

SHEPP = np.array([[1.0, 0.69, 0.92, 0.0, 0.0, 0.0],
                  [-0.8, 0.6624, 0.874, 0.0, -0.0184, 0.0],
                  [-0.2, 0.1100, 0.31, 0.22, 0.0, -math.pi / 10.0],
                  [-0.2, 0.1600, 0.41, -0.22, 0.0, math.pi / 10.0],
                  [0.1, 0.2100, 0.25, 0.0, 0.35, 0.0],
                  [0.1, 0.0460, 0.046, 0.0, 0.1, 0.0],
                  [0.1, 0.0460, 0.046, 0.0, -0.1, 0.0],
                  [0.1, 0.0460, 0.023, -0.08, -0.605, 0.0],
                  [0.1, 0.0230, 0.023, 0.0, -0.606, 0.0],
                  [0.1, 0.0230, 0.046, 0.06, -0.605, 0.0]])


# (name, n, image corners, #rays, #angles): arbitrary (t, theta) tables for withmesh
MESH_CASES = [
    ("mesh_n33_sorted", 33, DEF_IMG, 29, 23),
    ("mesh_n40_unsorted", 40, DEF_IMG, 37, 19),
    ("mesh_n24_fov", 24, (-0.7, 1.3, 1.2, -0.9), 30, 16),
]
# angles every mesh case also holds: the branch switch at pi/4, exact axes, wrap-around
SPECIAL_THETA = [0.0, math.pi / 4, math.pi / 2, 3 * math.pi / 4, math.pi, -math.pi / 2, 2 * math.pi]


# line-traversal (Bresenham / DDA) pair: (name, (lines, cols), R, A, image corners, sinogram
# corners); outputs from the reference's own iterator raft/raft_ddaiterator.c driven by
# oracle/ref_dda_harness.c (the reference's worker files need a BLAS this image lacks)
DDA_CASES = [
    ("dda_24x24_r30_a20", (24, 24), 30, 20, DEF_IMG, (0.0, 1.0, math.pi, -1.0)),
    ("dda_20x31_r25_a17_rect", (20, 31), 25, 17, (-0.7, 1.3, 1.2, -0.9), (-math.pi / 2, 1.1, math.pi / 2, -0.8)),
    ("dda_16x16_r21_a9_flip", (16, 16), 21, 9, (1.0, -1.0, -1.0, 1.0), (math.pi, -1.2, 0.0, 1.2)),
]


def mesh_tables(rng, R, A, sort):
    t = rng.uniform(-1.6, 1.6, R)           # some rays miss the image
    t[:3] = (-1.0, 0.0, 1.0)
    if sort:
        t.sort()
    theta = np.concatenate([rng.uniform(-4.0, 7.0, A - len(SPECIAL_THETA)), SPECIAL_THETA])
    return t, theta


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libraft_ref.so is missing: run make -C oracle ref")
    rng = np.random.default_rng(2024)
    out = {}
    for name, n, R, A, ic, sc in CASES:
        geo = dict(img_tl=ic[:2], img_br=ic[2:], sino_tl=sc[:2], sino_br=sc[2:])
        img = rng.random((n, n))
        sino = rng.random((R, A))
        out[name + "/img"] = img
        out[name + "/sino"] = sino
        out[name + "/geom"] = np.array([n, R, A] + list(ic) + list(sc), dtype=np.float64)
        out[name + "/radon"] = O.ref_radon(img, R, A, nthreads=1, **geo)
        out[name + "/backprojection"] = O.ref_backprojection(sino, n, nthreads=1, **geo)
    # Shepp-Logan rasterised by the reference, and its exact Radon (radon_fromdesc)
    out["shepp/desc"] = SHEPP
    out["shepp/phantom128"] = O.ref_phantom_fromdesc(SHEPP, 128)
    out["shepp/phantom33_fov"] = O.ref_phantom_fromdesc(SHEPP, 33, tl=(-0.9, 1.1), br=(1.2, -1.0))
    out["shepp/radon_fromdesc_96x60"] = O.ref_radon_fromdesc(SHEPP, 96, 60)
    out["shepp/radon_fromdesc_40x7_fov"] = O.ref_radon_fromdesc(SHEPP, 40, 7, tl=(0.2, 1.2), br=(2.9, -0.7))
    # arbitrary meshes (raft_radon_slantstack_withmesh) and single views
    for name, n, ic, R, A in MESH_CASES:
        img = rng.random((n, n))
        t, theta = mesh_tables(rng, R, A, sort=name.endswith("sorted") and "unsorted" not in name)
        out[name + "/img"] = img
        out[name + "/corners"] = np.array(ic, dtype=np.float64)
        out[name + "/t"] = t
        out[name + "/theta"] = theta
        out[name + "/radon"] = O.ref_radon_withmesh(img, t, theta, img_tl=ic[:2], img_br=ic[2:], nthreads=1)
        out[name + "/view"] = O.ref_radon_view(img, t, float(theta[0]), img_tl=ic[:2], img_br=ic[2:])
    for name, shape, R, A, ic, sc in DDA_CASES:
        geo = dict(img_tl=ic[:2], img_br=ic[2:], sino_tl=sc[:2], sino_br=sc[2:])
        img = rng.random(shape)
        sino = rng.random((R, A))
        out[name + "/img"] = img
        out[name + "/sino"] = sino
        out[name + "/geom"] = np.array([shape[0], shape[1], R, A] + list(ic) + list(sc), dtype=np.float64)
        out[name + "/radon"] = O.ref_radon_bresenham(img, R, A, **geo)
        out[name + "/transpose"] = O.ref_radon_transpose_bresenham(sino, shape, **geo)
        out[name + "/backprojection"] = O.ref_backprojection_bresenham(sino, shape, **geo)
    # a binary image file written by the reference's raft_image_write (raft/raft_image.c:135-151)
    io_img = rng.random((5, 7))
    out["io/values"] = io_img
    out["io/corners"] = np.array([-0.5, 2.0, 1.5, -1.0])
    rc = O.ref_image_write(os.path.join(HERE, "image_written_by_reference.raft"), io_img, (-0.5, 2.0), (1.5, -1.0))
    assert rc == 0
    path = os.path.join(HERE, "slantstack_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
