"""pyraft_amd -- MI355X-native slant-stack projector pair, FBP and EM/OSEM.

A drop-in for PyRAFT's slant-stack path (tacaswell/pyraft): the Python entry points
keep pyraft's names, signatures and (rays, angles) / (N, N) layouts and call the
libraft C ABI through ctypes; the compute runs in hand-written HIP kernels for
gfx950.  ``pyraft_amd.device`` (device-resident tensors) and ``pyraft_amd.autograd``
(differentiable projector) need PyTorch and are imported on demand.
"""
from .raftypes import *  # noqa: F401,F403
from .sinogram import *  # noqa: F401,F403
from .backprojection import *  # noqa: F401,F403
from .filters import *  # noqa: F401,F403
from .inversion import *  # noqa: F401,F403
from .iterative import *  # noqa: F401,F403
from .phantom import *  # noqa: F401,F403

__version__ = "0.1.0"
