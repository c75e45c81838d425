"""CPU: the 3D sweep kernel's exact skipping rules, modelled node by node in
tests/skipmodel/skip_model.c, never skip a node whose update would change it.

The model applies brick()'s decisions (per-node change bytes that only accumulate in the
first set with the seeded corners set, per-tile bytes, 6 face bytes per tile covering every
node of the face, whole-tile neighbour bytes below 8-plane tiles) and re-evaluates every
skipped node with the oracle's node operator; a sound skip finds the node's current time.
Sources sit on faces, edges, corners and inside; slow surface layers make head waves run
along the grid's first planes, which no pass in their direction updates.  The converged
grids and set counts must equal the oracle's, and a planted mistake in the node rule (the
z neighbour left out) must be caught.
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def skip_model(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("skipmodel") / "libskipmodel.so")
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off",
                    "-fno-fast-math", "-Wno-unused-function", "-o", so,
                    os.path.join(HERE, "skipmodel", "skip_model.c"), "-lm"], check=True)
    lib = ctypes.CDLL(so)
    f = lib.skip_model3d
    D, I = ctypes.c_double, ctypes.c_int
    f.argtypes = [ctypes.c_void_p, I, I, I, D, D, D, D, D, D, I, I, I, ctypes.c_void_p,
                  ctypes.POINTER(I), ctypes.POINTER(ctypes.c_longlong),
                  ctypes.POINTER(ctypes.c_longlong)]

    def run(slow, d, src, tile_a, plant=0, max_sets=100):
        slow = np.ascontiguousarray(slow, dtype=np.float64)
        nz, nx, ny = slow.shape
        tt = np.empty((nz + 1, nx + 1, ny + 1))
        k, ev, bad = I(), ctypes.c_longlong(), ctypes.c_longlong()
        f(slow.ctypes.data, nz, nx, ny, *d, *src, max_sets, tile_a, plant, tt.ctypes.data,
          ctypes.byref(k), ctypes.byref(ev), ctypes.byref(bad))
        return tt, k.value, ev.value, bad.value

    return run


def _case(seed):
    rng = np.random.default_rng(seed)
    shape = tuple(int(n) for n in rng.integers(9, 21, 3))
    p = rng.standard_normal(shape)
    for ax in range(3):
        p = (np.roll(p, 1, ax) + p + np.roll(p, -1, ax)) / 3.0
    v = 1500.0 * np.exp(rng.choice([0.3, 1.0]) * p / np.abs(p).max())
    if seed % 2 == 0:
        v[: rng.integers(1, 3)] *= rng.choice([0.35, 2.5])  # slow or fast surface layer
    d = tuple(float(h) for h in rng.uniform(6.0, 14.0, 3))
    ext = np.array(shape) * d
    src = rng.uniform(0.0, 1.0, 3) * ext
    where = seed % 4
    if where == 1:  # on a face
        src[rng.integers(3)] = 0.0
    elif where == 2:  # on an edge
        src[rng.choice(3, 2, replace=False)] = 0.0
    elif where == 3:  # in a corner (the far one)
        src = ext.copy()
    return 1.0 / v, d, src


@pytest.mark.parametrize("seed", range(16))
def test_skips_are_exact(skip_model, seed):
    slow, d, src = _case(seed)
    ref, _, _, kref = oracle.solve3d(slow, *d, src, nsweep=0)
    for tile_a in (4, 8, 16):
        tt, k, evals, bad = skip_model(slow, d, src, tile_a)
        assert bad == 0, (seed, tile_a)
        assert k == kref and np.array_equal(tt, ref), (seed, tile_a)
        assert evals < k * 8 * tt[1:, 1:, 1:].size  # the rules did skip work


def test_planted_rule_mistake_is_caught(skip_model):
    """Leaving the z neighbour out of the node rule skips nodes that would change; the
    re-evaluation counts them."""
    for seed in range(4):
        slow, d, src = _case(seed)
        assert skip_model(slow, d, src, 8, plant=1)[3] > 0, seed
