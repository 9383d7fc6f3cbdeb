"""pyraft_amd -- MI355X-native slant-stack projector pair, FBP and EM/OSEM.

A drop-in for PyRAFT's slant-stack path (tacaswell/pyraft): the Python entry points
keep pyraft's names, signatures and (rays, angles) / (N, N) layouts and call the
libraft C ABI through ctypes; the compute runs in hand-written HIP kernels for
gfx950.
"""
from .raftypes import *  # noqa: F401,F403
