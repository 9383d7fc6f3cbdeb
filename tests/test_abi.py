"""C ABI of libraft, checked without a GPU.

* the library loads and exports every function declared in include/*.h;
* the by-value structs have the reference's sizes;
* the plan logic (host build of the same FP64 code the device plan kernel runs)
  reproduces the reference predicate exactly, sample by sample;
* every index the HIP kernels form from the plan stays inside its buffer:
  Radon gather rows (fixed-point floor in [-1, N-1], covered by the zero pad) and
  the backprojection's LDS window (emulating the kernel's FP32 arithmetic);
* the host-side ramp kernel equals Re(ifft(fftshift(h))).
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEF_IC = (-1.0, 1.0, 1.0, -1.0)
DEF_SC = (0.0, 1.0, math.pi, -1.0)


@pytest.fixture(scope="module")
def lib():
    from pyraft_amd.raftypes import libraft
    return libraft


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*\**\s*([a-z_][a-z0-9_]*)\s*\(",
                             text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert "raft_radon_slantstack" in names and "raft_ss_em" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    from pyraft_amd.raftypes import EXPORTED_SYMBOLS
    assert set(EXPORTED_SYMBOLS) <= set(names)


def test_struct_sizes():
    from pyraft_amd.raftypes import RAFT_IMAGE, RAFT_MATRIX, RAFT_VECTOR
    assert ctypes.sizeof(RAFT_MATRIX) == 24
    assert ctypes.sizeof(RAFT_IMAGE) == 56
    assert ctypes.sizeof(RAFT_VECTOR) == 16


def host_plan(lib, n, R, A, ic=DEF_IC, sc=DEF_SC):
    b = np.zeros((A, R, 2), np.int32)
    fx = np.zeros((A, 3), np.int64)
    br = np.zeros(A, np.int32)
    rc = lib.raft_ss_plan_host_bounds(n, R, A, (ctypes.c_double * 4)(*ic), (ctypes.c_double * 4)(*sc),
                                      b.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                      fx.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                      br.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    assert rc == 0
    return b, fx, br


GEOMS = [
    (8, 8, 6, DEF_IC, DEF_SC), (17, 17, 45, DEF_IC, DEF_SC), (64, 64, 180, DEF_IC, DEF_SC),
    (64, 67, 45, DEF_IC, DEF_SC), (33, 40, 7, DEF_IC, DEF_SC), (128, 128, 4, DEF_IC, DEF_SC),
    (65, 64, 360, DEF_IC, DEF_SC), (200, 200, 100, DEF_IC, DEF_SC),
    (64, 64, 45, (-0.7, 1.3, 1.2, -0.9), (-math.pi / 2, 1.1, math.pi / 2, -0.8)),
    (64, 64, 45, DEF_IC, (0.3, 1.0, 2.5, -1.0)),
    (50, 31, 64, (1.0, 1.0, -1.0, -1.0), (math.pi, -1.0, 0.0, 1.0)),  # flipped axes
]


@pytest.mark.parametrize("geom", GEOMS, ids=lambda g: "n%d_R%d_A%d" % g[:3])
def test_plan_ranges_match_reference_predicate(lib, geom):
    n, R, A, ic, sc = geom
    b, fx, br = host_plan(lib, n, R, A, ic, sc)
    mask, coord = O.radon_sample_map(n, R, A, ic[:2], ic[2:], sc[:2], sc[2:])
    s = np.arange(n)
    irregular = b[..., 1] == -2
    inrange = (s >= b[..., 0][..., None]) & (s <= b[..., 1][..., None])
    assert not np.any((inrange != mask) & ~irregular[..., None])
    # irregular rays are exactly those whose valid set is not one interval
    for j, k in np.argwhere(irregular):
        idx = np.nonzero(mask[j, k])[0]
        assert idx.size and idx[-1] - idx[0] + 1 != idx.size
    # Q32.32 trajectory: within 1e-6 grid units of the reference coordinate and
    # floor in [-1, N-1] on every sample the reference accumulates
    k = np.arange(R)
    u = fx[:, 0][:, None, None] + k[None, :, None] * fx[:, 1][:, None, None] + s[None, None, :] * fx[:, 2][:, None, None]
    m = u >> 32
    assert np.all(np.abs(u / 2.0 ** 32 - coord)[mask] < 1e-6)
    assert m[mask].min() >= -1 and m[mask].max() <= n - 1


@pytest.mark.parametrize("n,A", [(2048, 1500), (1024, 720), (512, 720), (1536, 997), (768, 1800)])
def test_large_plan_index_ranges(lib, n, A):
    """Endpoints of every ray's range bound all its samples (u is linear in s)."""
    R = n
    b, fx, br = host_plan(lib, n, R, A)
    lo = b[..., 0].astype(np.int64)
    hi = b[..., 1].astype(np.int64)
    ok = hi >= lo
    k = np.arange(R)[None, :]
    base = fx[:, 0][:, None] + k * fx[:, 1][:, None]
    m = np.concatenate([((base + lo * fx[:, 2][:, None]) >> 32)[ok], ((base + hi * fx[:, 2][:, None]) >> 32)[ok]])
    assert m.min() >= -1 and m.max() <= n - 1
    assert np.all((lo >= 0) & (hi <= n - 1) | ~ok)


BP_ROWS = (4, 2, 8, 4, 2, 8)          # pixel rows per thread of each backprojection variant
BP_STAGE = (4096, 2048, 4096, 2048, 4096, 2048)


def _bp_window_extremes(lib, n, R, A, ic=DEF_IC, sc=DEF_SC, variant=0):
    """Emulate the kernel's FP32 window index for every pixel tile of variant
    `variant` (32 columns x 8*rows rows) at the tile corners; return the extremes
    and the window W the library stages."""
    Wv = (ctypes.c_int * len(BP_ROWS))()   # RAFT_SS_BP_NVARIANTS entries each
    Jv = (ctypes.c_int * len(BP_ROWS))()
    assert lib.raft_ss_plan_host_bp_staging(n, R, A, (ctypes.c_double * 4)(*ic), (ctypes.c_double * 4)(*sc),
                                            Wv, Jv) == 0
    W, J = Wv[variant], Jv[variant]
    if J == 0:
        return None
    assert 1 <= J <= 64 and J * W <= BP_STAGE[variant]
    rows, stage = ctypes.c_int(), ctypes.c_int()
    assert lib.raft_ss_bp_variant_shape(variant, ctypes.byref(rows), ctypes.byref(stage)) == 0
    assert (rows.value, stage.value) == (BP_ROWS[variant], BP_STAGE[variant])
    tile_h = 8 * rows.value
    xmin, ymin = ic[0], ic[3]
    dx = (ic[2] - ic[0]) / (n - 1)
    dy = (ic[1] - ic[3]) / (n - 1)
    tmin, dt = sc[3], (sc[1] - sc[3]) / (R - 1)
    th = sc[0] + np.arange(A) * ((sc[2] - sc[0]) / A)
    c = np.array([math.cos(t) for t in th])
    s = np.array([math.sin(t) for t in th])
    bpa, bpb, bpc = (xmin * c + ymin * s - tmin) / dt, (dx * c) / dt, (dy * s) / dt
    lo, hi = np.inf, -np.inf
    f32 = np.float32
    for c0 in range(0, n, 32):
        for r0 in range(0, n, tile_h):
            U = bpa + c0 * bpb + r0 * bpc
            umin = U + np.minimum(0, 31 * bpb) + np.minimum(0, (tile_h - 1) * bpc)
            base = np.floor(umin) - 1
            f0 = (U - base).astype(f32)
            for tx in (0, 31):
                t0 = (np.float64(tx) * bpb.astype(f32).astype(np.float64) + f0).astype(f32)
                for ty in (0, tile_h - 1):
                    u = (np.float64(ty) * bpc.astype(f32).astype(np.float64) + t0).astype(f32)
                    ii = np.floor(u)
                    lo, hi = min(lo, ii.min()), max(hi, (ii + 1).max())
    return lo, hi, W


@pytest.mark.parametrize("variant", range(len(BP_ROWS)))
@pytest.mark.parametrize("n,R,A", [(64, 64, 45), (65, 70, 37), (512, 512, 720), (2048, 2048, 1500), (100, 300, 50)])
def test_backprojection_window_in_bounds(lib, n, R, A, variant):
    """The kernel reads wrow[(int)u] and the pair's difference to ray (int)u + 1: both
    inside the staged window, and u >= 0 so truncation equals the floor."""
    r = _bp_window_extremes(lib, n, R, A, variant=variant)
    if r is None:
        pytest.skip("window does not fit this variant's stage")
    lo, hi, W = r
    assert lo >= 0 and hi <= W - 1, (lo, hi, W)


def _rt_window_extremes(lib, n, R, A, ic=DEF_IC, sc=DEF_SC, variant=3):
    """Emulate the pixel-driven transpose's FP32 window index floor(kappa - base) for the
    corner pixels of every tile of `variant` and every angle; return the extremes and the
    window W the library stages (None when the form does not apply)."""
    Wv = (ctypes.c_int * len(BP_ROWS))()
    Jv = (ctypes.c_int * len(BP_ROWS))()
    table = np.zeros((A, 5))
    assert lib.raft_ss_plan_host_rt_staging(n, R, A, (ctypes.c_double * 4)(*ic), (ctypes.c_double * 4)(*sc),
                                            Wv, Jv, table.ctypes.data) == 0
    W, J = Wv[variant], Jv[variant]
    if J == 0:
        return None
    # a stage of 2 STAGE floats holds J windows of W + 1 rays, W + 4 floats apart (with
    # slice pairs J / 2 angles of two interleaved windows: the same floats)
    assert 1 <= J <= 64 and J * (W + 4) <= 2 * BP_STAGE[variant]
    assert np.all(table[:, 4] >= 1.0)
    tile_h = 8 * BP_ROWS[variant]
    ig, a0, rb, br = table[:, 0], table[:, 1], table[:, 2], table[:, 3] != 0
    kb = np.where(br, -rb, ig)  # kappa per column (the kernel's expressions)
    kc = np.where(br, ig, -rb)  # kappa per row
    lo, hi = np.inf, -np.inf
    f32 = np.float32
    for c0 in range(0, n, 32):
        for r0 in range(0, n, tile_h):
            K = -a0 + c0 * kb + r0 * kc
            kmin = K + np.minimum(0, 31 * kb) + np.minimum(0, (tile_h - 1) * kc)
            base = np.floor(kmin) - 1
            f0 = (K - base).astype(f32)
            for tx in (0, 31):
                t0 = (np.float64(tx) * kb.astype(f32).astype(np.float64) + f0).astype(f32)
                for ty in (0, tile_h - 1):
                    k = (np.float64(ty) * kc.astype(f32).astype(np.float64) + t0).astype(f32)
                    ii = np.floor(k)
                    lo, hi = min(lo, ii.min()), max(hi, ii.max())
    return lo, hi, W


@pytest.mark.parametrize("variant", [3, 0, 1])
@pytest.mark.parametrize("n,R,A,ic,sc", [(64, 64, 45, DEF_IC, DEF_SC), (65, 61, 37, DEF_IC, DEF_SC),
                                         (512, 512, 720, DEF_IC, DEF_SC), (2048, 2048, 1500, DEF_IC, DEF_SC),
                                         (100, 40, 50, DEF_IC, DEF_SC),
                                         (48, 40, 91, (-0.7, 1.3, 1.2, -0.9), (-math.pi / 2, 1.1, math.pi / 2, -0.8))])
def test_transpose_window_in_bounds(lib, n, R, A, ic, sc, variant):
    """The pixel-driven transpose reads the staged rays floor(kappa - base) and the next
    one: both inside the window of W + 1 rays, and kappa - base >= 0 so truncation equals
    the floor."""
    r = _rt_window_extremes(lib, n, R, A, ic, sc, variant=variant)
    if r is None:
        pytest.skip("rays closer than a cell, or the window does not fit this variant")
    lo, hi, W = r
    assert lo >= 0 and hi <= W - 1, (lo, hi, W)


def test_transpose_form_needs_rays_a_cell_apart(lib):
    """Detectors with bins narrower than the pixels (rays closer than one cell) keep the
    line form: no pixel-driven staging."""
    Wv = (ctypes.c_int * len(BP_ROWS))()
    Jv = (ctypes.c_int * len(BP_ROWS))()
    assert lib.raft_ss_plan_host_rt_staging(64, 150, 30, (ctypes.c_double * 4)(*DEF_IC),
                                            (ctypes.c_double * 4)(*DEF_SC), Wv, Jv, None) == 0
    assert all(Jv[v] == 0 for v in range(len(BP_ROWS)))


@pytest.mark.parametrize("R", [64, 65, 256, 367])
def test_ramp_kernel_matches_numpy(lib, R):
    g = np.zeros(R, np.float32)
    assert lib.raft_ss_ramp_kernel(R, -1.0, 1.0, g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0
    ref = O.ramp_kernel(R)
    np.testing.assert_allclose(g, ref, rtol=2e-6, atol=1e-7 * np.abs(ref).max())


def test_invalid_geometry_reports_error(lib):
    b = np.zeros((4, 1, 2), np.int32)
    rc = lib.raft_ss_plan_host_bounds(1, 4, 4, (ctypes.c_double * 4)(*DEF_IC), (ctypes.c_double * 4)(*DEF_SC),
                                      b.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), None, None)
    assert rc == -1
    assert b"image side" in lib.raft_ss_last_error()
    rc = lib.raft_ss_plan_host_bounds(16, 16, 4, (ctypes.c_double * 4)(0.0, 1.0, 0.0, -1.0),
                                      (ctypes.c_double * 4)(*DEF_SC),
                                      np.zeros((4, 16, 2), np.int32).ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                      None, None)
    assert rc == -1 and b"zero-extent" in lib.raft_ss_last_error()
