"""CPU: why the library offers no hyperplane-parallel 2D sweep.

The reference's 2D set (fteikpy/_fteik/_fteik2d.py:190-287) interleaves the z passes of
every column inside each x half-set, so it is one serial dependency chain; the library
keeps that order (speculative 64-node blocks on one wave per source).  A parallel form
would run the four directional passes one after another, each over all columns (every
pass a node DAG whose anti-diagonals could update in parallel).  The oracle implements
that order (order=1).  Swept to convergence it reaches a different fixed point: FTeik's
admissibility tests make the fixed point depend on the update order.  On the benchmark's
2D models the relative difference from the reference's converged grid is 1e-5..1e-3, far
above the 1e-9 bar for converged traveltimes, so the mode is not offered.  This test pins
that measurement (the one model where both orders agree, the single-source 101 x 301
case, is recorded too).
"""
import numpy as np

import oracle
from fteikpy_amd import synthetic


def _drift(v, d, src):
    ref = oracle.solve2d(1.0 / v, *d, src, nsweep=0, order=0, nthreads=4)[0]
    par = oracle.solve2d(1.0 / v, *d, src, nsweep=0, order=1, nthreads=4)[0]
    return float((np.abs(par - ref) / np.abs(ref).clip(1e-300)).max())


def test_single_source_case_agrees():
    v, d, src, _ = synthetic.model_2d_gaussians()
    assert _drift(v, d, src[None, :]) == 0.0


def test_four_pass_order_leaves_the_reference_fixed_point():
    v, d, _, _ = synthetic.model_2d_gaussians()
    srcs = np.random.default_rng(3).random((64, 2)) * np.array(v.shape) * d
    worst = [_drift(v, d, srcs)]
    for v2, d2, s2 in synthetic.uneven_batch(n3=0)[:4]:
        worst.append(_drift(v2, d2, s2))
    print("max relative drift per model:", ["%.2e" % w for w in worst])
    assert min(worst) > 1e-9 and max(worst) > 1e-4
