"""FBP filters beyond lowpassino_fbp: the reference's C ramp filter raft_filter1d_ramp
(raft/raft_filter.cpp:118-213) with raft_fbp (raft/raft_fbp.c:5-24), and
lowpassino_reg_fbp (pyraft/filters/filters.py:73-111).

The reference's C filter needs FFTW, which this image lacks, so its oracle is a numpy
restatement (oracle.filter1d_ramp); these tests pin that restatement's FFT conventions
(zero padding to 2R - 1, the unnormalised inverse, which bins the ramp scales) with an
explicit DFT, and the library's Toeplitz-tap form of it against the restatement.  The
filter's definition itself is not pinned against reference outputs: none exist.
The numpy lowpassino filters are pinned by their eigenvalues: a cosine of bin k comes
out scaled by (G[k] + G[-k]) / 2, G = fftshift(h), which fixes the shift direction for
odd R.
"""
import ctypes
import math

import numpy as np
import pytest

from oracle import oracle as O

REL = 1e-5


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def brute_filter1d_ramp(x, ramp):
    """raft/raft_filter.cpp:70-92 for one column with explicit sums: Z = 2R - 1 point DFT
    of the zero-padded column, bins k <= R/2 (and their mirror Z - k) times ramp[k]/Z,
    the rest untouched, unnormalised inverse, first R samples.  `ramp` is given by the
    caller (hand-written values below), so the filter's definition is not re-derived."""
    R = len(x)
    Z = 2 * R - 1
    F = R // 2 + 1
    assert len(ramp) == F
    n = np.arange(Z)
    xz = np.zeros(Z)
    xz[:R] = x
    X = np.array([np.sum(xz * np.exp(-2j * np.pi * k * n / Z)) for k in range(Z)])
    H = np.ones(Z)
    for k in range(F):
        H[k] = ramp[k] / Z
        if k:
            H[Z - k] = ramp[k] / Z
    y = np.array([np.sum(H * X * np.exp(2j * np.pi * np.arange(Z) * m / Z)) for m in range(R)])
    return y.real


PI = math.pi
# raft/raft_filter.cpp:199-206 by hand for t in [-1, 1] (dt = 2 / (R - 1)):
#   F = R/2 + 1 values, delta = pi / (F - 1) / dt, ramp[i] = delta * i for i < int(F * cutoff)
HAND_RAMPS = {
    (7, 1.0): PI * np.array([0.0, 1.0, 2.0, 3.0]),                       # F = 4, dt = 1/3: delta = pi
    (8, 1.0): 7.0 * PI / 8.0 * np.array([0.0, 1.0, 2.0, 3.0, 4.0]),     # F = 5, dt = 2/7: delta = 7 pi / 8
    (8, 2.0): 7.0 * PI / 8.0 * np.array([0.0, 1.0, 2.0, 3.0, 4.0]),     # cutoff clamped to 1
    (12, 0.5): 11.0 * PI / 12.0 * np.array([0.0, 1.0, 2.0, 0, 0, 0, 0]),  # F = 7: int(3.5) = 3 values
    (9, 0.6): PI * np.array([0.0, 1.0, 2.0, 0.0, 0.0]),                # F = 5, dt = 1/4: 5 * 0.6 == 3.0
    (10, 0.0): np.zeros(6),                                             # cutoff 0: no bin kept
}


def test_filter1d_ramp_vector_by_hand():
    tl, br = (0.0, 1.0), (PI, -1.0)
    for (R, cutoff), want in HAND_RAMPS.items():
        got = O.filter1d_ramp_vector(R, cutoff, tl, br)
        np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-14, err_msg=str((R, cutoff)))
    # a detector span of [-2, 2]: dt = 4 / (R - 1), the slope halves
    np.testing.assert_allclose(O.filter1d_ramp_vector(7, 1.0, (0.0, 2.0), (PI, -2.0)),
                               PI / 2.0 * np.array([0.0, 1.0, 2.0, 3.0]), rtol=1e-14)


@pytest.mark.parametrize("R,cutoff", sorted(HAND_RAMPS))
def test_filter1d_ramp_restatement_matches_explicit_dft(R, cutoff):
    rng = np.random.default_rng(R)
    sino = rng.random((R, 3))
    tl, br = (0.0, 1.0), (PI, -1.0)
    got = O.filter1d_ramp(sino, cutoff, tl, br)
    for a in range(3):
        ref = brute_filter1d_ramp(sino[:, a], HAND_RAMPS[(R, cutoff)])
        assert np.abs(got[:, a] - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max()), a


@pytest.mark.parametrize("R", [7, 8, 10, 13])
def test_filter1d_ramp_cutoff_zero_closed_form(R):
    """cutoff = 0: bins 0..R/2 of the 2R - 1 point spectrum are zeroed, bins R/2+1..R-1
    (and their mirrors) pass unscaled, and the inverse is unnormalised (factor Z).  So
    the output is Z times the first R samples of the zero-padded column with its
    lowest R/2 + 1 frequencies (and mirrors) projected out.  Built from the real
    cosine/sine basis of the Z-point grid, not from an FFT."""
    rng = np.random.default_rng(100 + R)
    x = rng.random(R)
    Z = 2 * R - 1
    xz = np.zeros(Z)
    xz[:R] = x
    n = np.arange(Z)
    low = np.full(Z, xz.mean())                     # bin 0
    for k in range(1, R // 2 + 1):                  # bins 1..R/2 with their mirrors
        c, s = np.cos(2 * PI * k * n / Z), np.sin(2 * PI * k * n / Z)
        low += (2.0 / Z) * (c * np.dot(xz, c) + s * np.dot(xz, s))
    want = Z * (xz - low)[:R]
    got = O.filter1d_ramp(x[:, None], 0.0, (0.0, 1.0), (PI, -1.0))[:, 0]
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-10 * np.abs(want).max())
    # the reference's misreadings this separates: filtering every bin gives 0, and a
    # normalised inverse gives want / Z
    assert np.abs(want).max() > 0.1


@pytest.fixture(scope="module")
def lib():
    from pyraft_amd.raftypes import libraft
    return libraft


@pytest.mark.parametrize("R,cutoff", [(8, 1.0), (33, 0.7), (64, 1.0), (101, 0.25)])
def test_toeplitz_taps_reproduce_restatement(lib, R, cutoff):
    dt = -2.0 / (R - 1)
    taps = np.zeros(2 * R, np.float32)
    assert lib.raft_ss_filter1d_ramp_taps(R, cutoff, dt, taps.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0
    x = np.random.default_rng(R).random((R, 4))
    y = np.zeros_like(x)
    for n in range(R):
        for m in range(R):
            y[n] += x[m] * taps[n - m + R]
    ref = O.filter1d_ramp(x, cutoff, (0.0, 1.0), (math.pi, -1.0))
    assert rel_l2(y, ref) <= 1e-6


@pytest.mark.parametrize("R", [16, 17, 64, 65])
@pytest.mark.parametrize("kind,L", [(0, 0.0), (1, 1e-3), (1, 0.05)])
def test_numpy_filters_eigenvalues_and_kernels(lib, R, kind, L):
    """Cosines of bin k are eigenvectors with eigenvalue (G[k] + G[R-k]) / 2; the
    library's FP32 kernel convolves to the oracle."""
    tl, br = (0.0, 1.0), (math.pi, -1.0)
    t = np.linspace(br[1], tl[1], R)
    dt = (t[R - 1] - t[0]) / (R - 1)
    wc = 1 / (2 * dt)
    w = np.linspace(-wc, wc, R)
    h = np.abs(w) * np.cos(np.pi * w * dt) if kind == 0 else np.abs(w) / (1.0 + L * np.abs(w) ** 2)
    G = np.roll(h, R // 2)
    n = np.arange(R)
    for k in (0, 1, 3, R // 2):
        x = np.cos(2 * np.pi * k * n / R)[:, None]
        out = O.lowpassino_fbp(x, tl, br) if kind == 0 else O.lowpassino_reg_fbp(x, L, tl, br)
        assert np.allclose(out, 0.5 * (G[k] + G[(R - k) % R]) * x, atol=1e-9 * np.abs(G).max()), k
    g = np.zeros(R, np.float32)
    assert lib.raft_ss_filter_kernel(R, br[1], tl[1], kind, L, g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0
    x = np.random.default_rng(R).random((R, 3))
    y = np.stack([np.array([np.sum(x[:, a] * g[(m - n) % R]) for m in range(R)]) for a in range(3)], axis=1)
    ref = O.lowpassino_fbp(x, tl, br) if kind == 0 else O.lowpassino_reg_fbp(x, L, tl, br)
    assert rel_l2(y, ref) <= 1e-6


def test_filter_argument_checks(lib):
    taps = np.zeros(4, np.float32)
    p = taps.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    assert lib.raft_ss_filter1d_ramp_taps(1, 1.0, 0.1, p) < 0
    assert lib.raft_ss_filter1d_ramp_taps(2, 1.0, 0.0, p) < 0
    assert lib.raft_ss_filter_kernel(8, -1.0, 1.0, 7, 0.0, p) < 0
    import pyraft_amd as pr
    with pytest.raises(TypeError):
        pr.lowpassino_reg_fbp(pr.image((8, 4)))


# ---------------------------------------------------------------- GPU


@pytest.mark.gpu
def test_gpu_filter1d_ramp_and_fbp():
    import torch
    assert torch.cuda.is_available()
    import pyraft_amd as pr
    from pyraft_amd.raftypes import RAFT_BRESENHAM, RAFT_SLANTSTACK, check_legacy, libraft, make_RAFT_IMAGE
    R, A, n = 96, 60, 64
    rng = np.random.default_rng(2)
    base = rng.random((R, A))
    tl, br = (0.0, 1.0), (math.pi, -1.0)
    for cutoff in (1.0, 0.6):
        print("filter1d_ramp cutoff", cutoff, flush=True)
        s = base.copy()
        libraft.raft_filter1d_ramp(ctypes.c_double(cutoff), make_RAFT_IMAGE(s, tl, br), 4)
        check_legacy("raft_filter1d_ramp")
        assert rel_l2(s, O.filter1d_ramp(base, cutoff, tl, br)) <= REL
    filt = O.filter1d_ramp(base, 1.0, tl, br)
    for method in (RAFT_SLANTSTACK, RAFT_BRESENHAM):
        print("raft_fbp method", method, flush=True)
        s = base.copy()
        img = np.zeros((n, n))
        libraft.raft_fbp(ctypes.c_double(1.0), make_RAFT_IMAGE(s, tl, br), make_RAFT_IMAGE(img), method, 4)
        check_legacy("raft_fbp")
        assert rel_l2(s, filt) <= REL                   # filtered in place, as the reference does
        if method == RAFT_SLANTSTACK:
            ref = O.backprojection(filt, n, sino_tl=tl, sino_br=br)
        else:
            ref = O.backprojection_bresenham(filt, (n, n), sino_tl=tl, sino_br=br)
        ok = np.isfinite(ref)
        assert np.array_equal(ok, np.isfinite(img)) and rel_l2(img[ok], ref[ok]) <= REL
    # pyraft.lowpassino_reg_fbp
    y = pr.image(base, x_extent=(0.0, math.pi))
    for L in (1e-4, 1e-2):
        out = pr.lowpassino_reg_fbp(y, L)
        assert out.top_left == y.top_left and rel_l2(out, O.lowpassino_reg_fbp(base, L, tl, br)) <= REL


@pytest.mark.gpu
@pytest.mark.parametrize("R,nrows", [(2, 3), (7, 5), (96, 1), (333, 9), (1024, 6), (2048, 3), (2049, 2)])
def test_gpu_row_filter_forms(R, nrows):
    """The device row filter (FFT convolution up to R = 2048, the direct sum above) against
    an FP64 direct evaluation of the same taps: circular (lowpassino kernels) and Toeplitz
    (raft_filter1d_ramp), odd sizes and odd row counts (a block holding one row)."""
    import torch
    assert torch.cuda.is_available()
    from pyraft_amd.device import current_stream_handle, filter_rows
    from pyraft_amd.raftypes import check, libraft
    print("row filter", R, nrows, flush=True)
    rng = np.random.default_rng(R)
    x = rng.random((nrows, R)).astype(np.float32)
    g = np.zeros(R, np.float32)
    assert libraft.raft_ss_filter_kernel(R, -1.0, 1.0, 0, 0.0, g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0
    got = filter_rows(torch.from_numpy(x).cuda(), torch.from_numpy(g).cuda()).cpu().numpy()
    n = np.arange(R)
    C = g.astype(np.float64)[(n[None, :] - n[:, None]) % R]        # C[m][n] = g[(n - m) mod R]
    assert rel_l2(got, x.astype(np.float64) @ C) <= REL
    taps = np.zeros(2 * R, np.float32)
    assert libraft.raft_ss_filter1d_ramp_taps(R, 0.8, -2.0 / (R - 1),
                                              taps.ctypes.data_as(ctypes.POINTER(ctypes.c_float))) == 0
    xd, td = torch.from_numpy(x).cuda(), torch.from_numpy(taps).cuda()
    out = torch.empty_like(xd)
    check(libraft.raft_ss_filter_rows_toeplitz(xd.data_ptr(), out.data_ptr(), td.data_ptr(), R, nrows,
                                               current_stream_handle(xd.device)), "raft_ss_filter_rows_toeplitz")
    T = taps.astype(np.float64)[(n[None, :] - n[:, None]) + R]       # T[m][n] = taps[n - m + R]
    assert rel_l2(out.cpu().numpy(), x.astype(np.float64) @ T) <= REL
