"""Thread-count knob of fteikpy/_helpers.py:4-27.

The reference sizes numba's CPU thread pool; here the solvers run on the GPU, so the
value only sets torch's intra-op threads used by host-side array work.
"""
import torch


def get_num_threads():
    """Current number of host threads."""
    return torch.get_num_threads()


def set_num_threads(n):
    """Set the number of host threads."""
    torch.set_num_threads(int(n))
