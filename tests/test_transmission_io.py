"""Transmission EM (raft_tr_em_gpu, raft/raft_transmission_em.cu:16-66) and the binary
raft_matrix / raft_image files (raft/raft_matrix.c:102-136, 261-287; raft/raft_image.c).

CPU:
* files written by the reference's raft_image_write (tests/golden fixture, and live
  through oracle/_ref where it is built) read back bit for bit, and files this library
  writes read bit for bit with the reference's reader; truncated files give an empty
  result / an error, as the reference's reader does;
* the oracle's transmission EM keeps x* fixed when raw = flat exp(-R x*) (the update
  ratio is then 1), and stays positive.

GPU: raft_tr_em_gpu, raft_ss_tr_em (several slices per launch) and
pyraft.em_transmission_gpu against the FP64 oracle (rel-L2 <= 1e-5).
"""
import ctypes
import os

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden", "slantstack_golden.npz")
REF_FILE = os.path.join(HERE, "golden", "image_written_by_reference.raft")
REL = 1e-5


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# ---------------------------------------------------------------- binary files


def test_reads_file_written_by_reference():
    import pyraft_amd as pr
    g = np.load(GOLDEN)
    im = pr.image_read(REF_FILE)
    assert np.array_equal(np.asarray(im), g["io/values"])
    assert im.top_left == tuple(g["io/corners"][:2]) and im.bottom_right == tuple(g["io/corners"][2:])


@pytest.mark.skipif(not O.ref_available(), reason="reference checker oracle/_ref not built here")
def test_files_interchange_with_reference(tmp_path):
    import pyraft_amd as pr
    rng = np.random.default_rng(1)
    a = pr.image(rng.random((9, 4)), top_left=(0.25, 3.0), bottom_right=(2.0, -1.5))
    p = tmp_path / "ours.raft"
    pr.image_write(a, str(p))
    vals, corners = O.ref_image_read(p)
    assert np.array_equal(vals, a) and corners == (0.25, 3.0, 2.0, -1.5)
    # a strided view is written in row order
    b = np.asarray(rng.random((6, 10)))[::2, ::3]
    q = tmp_path / "ref.raft"
    assert O.ref_image_write(q, np.ascontiguousarray(b), (-1.0, 1.0), (1.0, -1.0)) == 0
    pr.image_write(pr.image(b), str(tmp_path / "ours2.raft"))
    assert (tmp_path / "ours2.raft").read_bytes() == q.read_bytes()


def test_truncated_files(tmp_path):
    import pyraft_amd as pr
    from pyraft_amd.raftypes import _libc, libraft
    a = pr.image(np.arange(6.0).reshape(2, 3))
    p = tmp_path / "a.raft"
    pr.image_write(a, str(p))
    data = p.read_bytes()
    assert len(data) == 8 + 6 * 8 + 4 * 8
    for cut in (4, 8 + 5 * 8, len(data) - 8):
        q = tmp_path / ("cut%d.raft" % cut)
        q.write_bytes(data[:cut])
        with pytest.raises(OSError):
            pr.image_read(str(q))
    # the matrix reader alone: a complete matrix, then nothing for a short one
    fp = _libc.fopen(str(p).encode(), b"rb")
    m = libraft.raft_matrix_read(fp)
    _libc.fclose(fp)
    assert (m.lines, m.columns, m.line_stride, m.column_stride) == (2, 3, 1, 3)
    assert np.ctypeslib.as_array(m.p_data, shape=(6,)).tolist() == list(range(6))
    libraft.raft_matrix_destroy(ctypes.byref(m))
    assert not m.p_data


# ---------------------------------------------------------------- transmission EM (oracle)


def test_oracle_tr_em_fixed_point_and_positivity():
    n, R, A = 32, 36, 30
    rng = np.random.default_rng(3)
    xs = 0.2 + 0.5 * rng.random((n, n))
    flat = 1000.0 * (0.5 + rng.random((R, A)))
    raw = flat * np.exp(-O.radon(xs, R, A))
    x = O.tr_em(raw, flat, n, 3, x0=xs)
    w = O.backprojection(raw, n)
    assert np.allclose(x[w != 0], xs[w != 0], rtol=1e-10)
    y = O.tr_em(raw, flat, n, 4)
    assert np.all(y > 0) and np.all(np.isfinite(y))


# ---------------------------------------------------------------- GPU


def _case(n, R, A, seed):
    rng = np.random.default_rng(seed)
    xs = 0.3 * rng.random((n, n))
    flat = 500.0 * (0.8 + 0.4 * rng.random((R, A)))
    raw = rng.poisson(flat * np.exp(-O.radon(xs, R, A))).astype(np.float64)
    return raw, flat


@pytest.mark.gpu
def test_gpu_tr_em_dropin_and_wrapper():
    import torch
    assert torch.cuda.is_available()
    import pyraft_amd as pr
    from pyraft_amd.raftypes import check_legacy, libraft
    n, R, A, niter = 64, 64, 45, 4
    raw, flat = _case(n, R, A, 7)
    ref = O.tr_em(raw, flat, n, niter)
    out = np.zeros(n * n, np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    r32, f32 = np.ascontiguousarray(raw, np.float32), np.ascontiguousarray(flat, np.float32)
    libraft.raft_tr_em_gpu(out.ctypes.data_as(fp), r32.ctypes.data_as(fp), f32.ctypes.data_as(fp), n, R, A, niter)
    check_legacy("raft_tr_em_gpu")
    assert rel_l2(out.reshape(n, n), ref) <= REL
    x = pr.em_transmission_gpu(raw, flat, np.zeros_like(raw), n, niter)
    assert x.dtype == np.float64 and rel_l2(x, ref) <= REL


@pytest.mark.gpu
def test_gpu_tr_em_slices():
    import torch
    from pyraft_amd.device import Plan
    from pyraft_amd.raftypes import check, libraft
    n, R, A, niter = 96, 100, 73, 3
    cases = [_case(n, R, A, s) for s in (11, 12, 13)]
    raw = torch.from_numpy(np.stack([c[0] for c in cases]).astype(np.float32)).cuda()
    flat = torch.from_numpy(np.stack([c[1] for c in cases]).astype(np.float32)).cuda()
    x = torch.ones((3, n, n), device="cuda")
    p = Plan(n, R, A)
    check(libraft.raft_ss_tr_em(p.handle, raw.data_ptr(), flat.data_ptr(), x.data_ptr(), niter, 3,
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "raft_ss_tr_em")
    got = x.cpu().numpy()
    for i, (r, f) in enumerate(cases):
        assert rel_l2(got[i], O.tr_em(r, f, n, niter)) <= REL, i
