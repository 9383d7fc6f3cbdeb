"""Host-side logic of the pyraft-compatible API (no GPU needed)."""
import math
import pickle
import subprocess
import sys
import os

import numpy as np
import pytest

import pyraft_amd as pr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_image_defaults_and_extents():
    im = pr.image((4, 6))
    assert im.dtype == np.float64 and np.all(im == 0)
    assert im.top_left == (-1.0, 1.0) and im.bottom_right == (1.0, -1.0)
    assert im.sampling_distances == (2.0 / 5, -2.0 / 3)
    s = pr.image((5, 3), x_extent=(0.0, math.pi))
    assert s.top_left == (0.0, 1.0) and s.bottom_right == (math.pi, -1.0)
    assert s.extent() == (0.0, math.pi, -1.0, 1.0)
    with pytest.raises(TypeError):
        pr.image((2, 2), top_left=(0, 1), x_extent=(0, 1))
    copy = pr.image(s)
    assert copy.top_left == s.top_left and copy is not s


def test_image_pickle_keeps_corners():
    im = pr.image(np.arange(6.0).reshape(2, 3), top_left=(0.5, 2.0), bottom_right=(1.5, -2.0))
    back = pickle.loads(pickle.dumps(im))
    assert back.top_left == (0.5, 2.0) and back.bottom_right == (1.5, -2.0)
    assert np.array_equal(back, im)


def test_radon_method_members_are_tuples():
    assert pr.radon_method.SLANTSTACK == (2,)
    assert pr.radon_method.BRESENHAM == (1,)


def test_plans_and_unsupported_methods():
    assert pr.backprojection_plan((16, 8), (16, 16), method=pr.radon_method.SLANTSTACK) == \
        (pr.radon_method.SLANTSTACK, (16, 16), 2)
    assert pr.backprojection_plan((16, 8), (16, 16), method=pr.radon_method.BRESENHAM) == \
        (pr.radon_method.BRESENHAM, (16, 16), 2)
    with pytest.raises(NotImplementedError):
        pr.backprojection_plan((16, 8), (16, 16))  # BST default
    with pytest.raises(TypeError):
        pr.backprojection_plan((16, 8), (16, 16), method=(99,))
    with pytest.raises(TypeError):
        pr.radon(pr.image((8, 8)), (8, 4), method=pr.radon_method.BST)
    with pytest.raises(TypeError):
        pr.radon_transpose(np.zeros((8, 4)), (8, 8), method=pr.radon_method.BST)
    # defaults as in pyraft: the line-traversal pair
    assert pr.radon.__defaults__[1] == pr.radon_method.BRESENHAM
    assert pr.radon_transpose.__defaults__[1] == pr.radon_method.BRESENHAM
    assert pr.make_em.__defaults__[:2] == (pr.radon_method.BRESENHAM, pr.radon_method.BRESENHAM)


def test_operators_fail_loudly_without_a_device():
    """No CPU fallback: with no GPU the drop-in symbols report a HIP error."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(pr.RaftError):
        pr.radon(pr.image(np.ones((8, 8))), (8, 4))
    with pytest.raises(pr.RaftError):
        pr.radon(pr.image(np.ones((8, 8))), (8, 4), method=pr.radon_method.SLANTSTACK)


def test_make_raft_matrix_strides_and_dtype():
    a = np.arange(12.0).reshape(3, 4)
    m = pr.make_RAFT_MATRIX(a)
    assert (m.lines, m.columns, m.line_stride, m.column_stride) == (3, 4, 1, 4)
    t = pr.make_RAFT_MATRIX(a.T)
    assert (t.lines, t.columns, t.line_stride, t.column_stride) == (4, 3, 4, 1)
    with pytest.raises(TypeError):
        pr.make_RAFT_MATRIX(a.astype(np.float32))


def test_random_ellipses_are_seeded_and_in_range():
    d1 = pr.random_ellipses(32, 2)
    d2 = pr.random_ellipses(32, 2)
    assert np.array_equal(d1, d2) and d1.shape == (32, 6)
    assert np.all((d1[:, 1] >= 0.02) & (d1[:, 1] <= 0.4) & (d1[:, 2] >= 0.02) & (d1[:, 2] <= 0.4))
    assert np.all(np.hypot(d1[:, 3], d1[:, 4]) <= 1.0)
    assert np.all((d1[:, 0] >= 0) & (d1[:, 0] <= 1))


def test_bench_refuses_more_than_one_gpu():
    proc = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                          capture_output=True, text=True, timeout=120)
    assert proc.returncode == 2
    assert "one GPU" in proc.stderr


def test_graft_entry_points_exist():
    import importlib.util
    spec = importlib.util.spec_from_file_location("ge", os.path.join(ROOT, "__graft_entry__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    assert callable(mod.build) and callable(mod.smoke)
