"""raft_ss_radon_noalloc (include/raft_noalloc.h): the slant-stack Radon with its
workspace in the library's static device array, called through the C ABI and compared
bit for bit with raft_ss_radon_range, which allocates its workspace per call.

Two calls in a row share the array: a 2048 x 2048 slice over all 1500 angles (the
largest workspace the array holds), then a smaller geometry over a few angles (a
divided launch with slab partial sums, two slices), so the second call runs on the
first one's leftovers.  A call needing more than the array holds is refused.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(fn, plan, x, out, angle_major, j0, j1, nslices):
    import torch
    from pyraft_amd.device import current_stream_handle
    return fn(plan.handle, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()), int(angle_major),
              int(j0), int(j1), int(nslices), current_stream_handle(x.device))


def test_radon_noalloc_matches_allocating_entry():
    import torch
    from pyraft_amd.device import Plan, phantom
    from pyraft_amd.phantom import random_ellipses
    from pyraft_amd.raftypes import libraft
    assert torch.cuda.is_available()
    noalloc = libraft.raft_ss_radon_noalloc
    alloc = libraft.raft_ss_radon_range
    cases = [
        # (n, R, A, angle range, slices, angle-major)
        (2048, 2048, 1500, (0, 1500), 1, True),
        (300, 310, 97, (5, 50), 2, False),
    ]
    for n, R, A, (j0, j1), ns, am in cases:
        print("radon_noalloc n=%d R=%d A=%d angles [%d, %d) slices %d" % (n, R, A, j0, j1, ns), flush=True)
        plan = Plan(n, R, A)
        x = torch.stack([phantom(random_ellipses(16, 40 + i), n) for i in range(ns)]).contiguous()
        shape = (ns, A, R) if am else (ns, R, A)
        got = torch.full(shape, -1.0, device="cuda")
        ref = torch.full(shape, -1.0, device="cuda")
        # the allocating entry first: it makes the plan's launch-shape choice, which the
        # allocation-free entry then reuses
        assert _call(alloc, plan, x, ref, am, j0, j1, ns) == 0, libraft.raft_ss_last_error()
        assert _call(noalloc, plan, x, got, am, j0, j1, ns) == 0, libraft.raft_ss_last_error()
        torch.cuda.synchronize()
        assert torch.equal(got, ref)
        cols = got[:, j0:j1] if am else got[..., j0:j1]
        assert float(cols.abs().sum()) > 0
    # a workspace beyond the array: refused with a message, nothing launched
    plan = Plan(2048, 2048, 16)
    x = torch.zeros((2, 2048, 2048), device="cuda")
    out = torch.full((2, 16, 2048), 7.0, device="cuda")
    rc = _call(noalloc, plan, x, out, True, 0, 16, 2)
    torch.cuda.synchronize()
    assert rc != 0
    assert b"static scratch" in libraft.raft_ss_last_error()
    assert bool((out == 7.0).all())
