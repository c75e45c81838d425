"""Thread-count knob, fteikpy/_helpers.py:4-27.

The reference sizes numba's CPU thread pool (numba.set_num_threads refuses counts outside
1..the pool size).  Here every solve, interpolation and ray runs on the GPU; the knob sizes
the host threads torch uses for the host-side array work around them (the ray lists,
meshio arrays, host copies), with the same accepted range: 1 .. the host's logical CPUs
available to this process.
"""
import os

import torch


def _host_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # platforms without affinity masks
        return os.cpu_count() or 1


def get_num_threads():
    """Current number of host threads."""
    return torch.get_num_threads()


def set_num_threads(n):
    """Set the number of host threads (1 .. available logical CPUs, else ValueError)."""
    if isinstance(n, bool) or int(n) != n:
        raise TypeError("The number of threads specified must be an integer")
    n = int(n)
    limit = _host_threads()
    if not 1 <= n <= limit:
        raise ValueError(f"The number of threads must be between 1 and {limit}")
    torch.set_num_threads(n)
