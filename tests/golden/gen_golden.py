"""Write golden input/output vectors from the reference fteikpy 2.4.0 (run in this
container only; the reference never travels with the repository).

The reference is run in pure-Python mode, the mode its own CI uses
(``NUMBA_DISABLE_JIT=1``, .github/workflows/ci.yml:29).  numba is not importable here, so
a minimal module is registered in ``sys.modules`` in its place: ``jit`` returns the
function unchanged (what NUMBA_DISABLE_JIT does) and ``prange`` is ``range``.

The cases are the ones the reference's own tests leave open (heterogeneous media,
non-unit spacing, non-zero origin, off-grid sources, edge queries, rays in heterogeneous
media); the reference's own test values are checked directly in tests/test_oracle.py.

Usage:  python tests/golden/gen_golden.py [/path/to/reference]
Output: tests/golden/golden_3d.npz, tests/golden/golden_2d.npz
"""
import os
import sys
import time
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _install_pure_python_numba():
    mod = types.ModuleType("numba")

    def jit(*args, **kwargs):
        if args and callable(args[0]):
            return args[0]
        return lambda f: f

    mod.jit = jit
    mod.njit = jit
    mod.prange = range
    state = {"n": 1}
    mod.get_num_threads = lambda: state["n"]
    mod.set_num_threads = lambda n: state.__setitem__("n", n)
    sys.modules["numba"] = mod


def smooth_field(rng, shape, passes):
    """N(0,1) noise smoothed by repeated 3-point averaging along every axis, scaled to
    max |p| = 1."""
    p = rng.standard_normal(shape)
    for _ in range(passes):
        for ax in range(len(shape)):
            p = (np.roll(p, 1, ax) + p + np.roll(p, -1, ax)) / 3.0
    return p / np.abs(p).max()


def model3d(rng, shape, d):
    """Rough high-contrast medium (velocity ratio up to ~10) with a fast slab: first
    arrivals bend and reverse direction, so convergence needs several sweep sets."""
    nz, nx, ny = shape
    v = 1500.0 * np.exp(1.1 * smooth_field(rng, shape, 1))
    v[nz // 3: nz // 3 + 2, : 2 * nx // 3, :] = 6000.0
    return v


def model2d(rng, shape, d):
    """Rough high-contrast 2D medium with a fast layer and a slow wall."""
    nz, nx = shape
    v = 1500.0 * np.exp(1.1 * smooth_field(rng, shape, 1))
    v[nz // 2: nz // 2 + 2, :] = 6000.0
    v[:, 2 * nx // 3] = 500.0
    return v


def converged_nsweep(solve, cap=16):
    """Smallest K such that set K changes no node (K counts that last set)."""
    prev = solve(1)
    for k in range(2, cap + 1):
        cur = solve(k)
        if np.array_equal(cur, prev):
            return k
        prev = cur
    raise RuntimeError("no convergence within cap")


def queries(rng, origin, d, n_nodes, nq):
    """Random queries plus node-exact edge queries and out-of-bound points."""
    lo = origin
    hi = origin + d * (np.asarray(n_nodes) - 1)
    q = lo + rng.random((nq, len(lo))) * (hi - lo)
    corners = np.array(np.meshgrid(*[[a, b] for a, b in zip(lo, hi)], indexing="ij")).reshape(len(lo), -1).T
    edge = q[: len(lo)].copy()
    for a in range(len(lo)):
        edge[a, a] = hi[a]  # exactly on the last node plane: ghost-node branch
    out = np.array([lo - 1.0, hi + 1.0])
    return np.vstack([q, corners, edge, out])


def main(ref):
    _install_pure_python_numba()
    sys.path.insert(0, ref)
    import fteikpy  # noqa: E402

    rng = np.random.default_rng(20261203)
    out3 = {}
    t0 = time.time()

    # ---------------- 3D ----------------
    shape3 = (12, 14, 16)
    d3 = np.array([10.0, 12.5, 8.0])
    origin3 = np.array([-5.0, 100.0, 2.5])
    v3 = model3d(rng, shape3, d3)
    out3["v"] = v3
    out3["gridsize"] = d3
    out3["origin"] = origin3
    eik3 = fteikpy.Eikonal3D(v3, d3, origin3)
    ext = d3 * np.array(shape3)
    srcs3 = np.array([
        origin3 + np.array([23.7, 61.3, 40.9]),  # off-grid interior
        origin3 + np.array([30.0, 50.0, 40.0]),  # on a node
        origin3 + np.array([ext[0], 0.3 * ext[1], ext[2]]),  # on the far faces (edge nudge)
    ])
    out3["sources"] = srcs3
    for s, src in enumerate(srcs3):
        kconv = converged_nsweep(lambda k: eik3.solve(src, nsweep=k).grid)
        out3[f"s{s}_kconv"] = np.array(kconv)
        for ns in (1, 2, 3, kconv):
            tt = eik3.solve(src, nsweep=ns, return_gradient=True)
            tag = f"s{s}_n{ns}" if ns != kconv else f"s{s}_conv"
            out3[tag + "_tt"] = tt.grid
            out3[tag + "_grad"] = np.stack([g.grid for g in tt.gradient], axis=-1)
            out3[tag + "_vzero"] = np.array(tt._vzero)
        tt = eik3.solve(src, nsweep=kconv, return_gradient=True)
        q = queries(rng, origin3, d3, tt.shape, 40)
        out3[f"s{s}_q"] = q
        out3[f"s{s}_vinterp"] = tt(q)
        out3[f"s{s}_gz_interp"] = tt.gradient[0](q)
        # rays from in-bound random receivers, both modes
        rec = q[:6]
        for hg in (False, True):
            rays = tt.raytrace(rec, stepsize=None, max_step=None, honor_grid=hg)
            lens = np.array([len(r) for r in rays])
            out3[f"s{s}_rays_hg{int(hg)}_len"] = lens
            out3[f"s{s}_rays_hg{int(hg)}"] = np.concatenate(rays)
        out3[f"s{s}_rec"] = rec
    # velocity interpolation on the model grid
    qv = queries(rng, origin3, d3, shape3, 30)
    out3["vq"] = qv
    out3["v_interp"] = eik3(qv)
    # batch solve (list return path)
    tts = eik3.solve(srcs3[:2], nsweep=2)
    out3["batch_tt"] = np.stack([t.grid for t in tts])
    print(f"3D done in {time.time() - t0:.1f} s", flush=True)

    # ---------------- 2D ----------------
    out2 = {}
    t0 = time.time()
    shape2 = (21, 31)
    d2 = np.array([10.0, 7.5])
    origin2 = np.array([3.0, -20.0])
    v2 = model2d(rng, shape2, d2)
    out2["v"] = v2
    out2["gridsize"] = d2
    out2["origin"] = origin2
    eik2 = fteikpy.Eikonal2D(v2, d2, origin2)
    ext2 = d2 * np.array(shape2)
    srcs2 = np.array([
        origin2 + np.array([12.3, 101.37]),  # off-grid, spherical box touching the top
        origin2 + np.array([100.0, 75.0]),  # on a node (iflag 1)
        origin2 + np.array([55.0, 113.9]),  # on a z node, off in x (iflag 2, one coordinate rounded)
        origin2 + np.array([ext2[0], 0.61 * ext2[1]]),  # bottom face
    ])
    out2["sources"] = srcs2
    for s, src in enumerate(srcs2):
        kconv = converged_nsweep(lambda k: eik2.solve(src, nsweep=k).grid)
        out2[f"s{s}_kconv"] = np.array(kconv)
        for ns in (1, 2, 3, kconv):
            tt = eik2.solve(src, nsweep=ns, return_gradient=True)
            tag = f"s{s}_n{ns}" if ns != kconv else f"s{s}_conv"
            out2[tag + "_tt"] = tt.grid
            out2[tag + "_grad"] = np.stack([g.grid for g in tt.gradient], axis=-1)
            out2[tag + "_vzero"] = np.array(tt._vzero)
        tt = eik2.solve(src, nsweep=kconv, return_gradient=True)
        q = queries(rng, origin2, d2, tt.shape, 40)
        out2[f"s{s}_q"] = q
        out2[f"s{s}_vinterp"] = tt(q)
        out2[f"s{s}_gz_interp"] = tt.gradient[0](q)
        rec = q[:6]
        for hg in (False, True):
            rays = tt.raytrace(rec, stepsize=None, max_step=None, honor_grid=hg)
            out2[f"s{s}_rays_hg{int(hg)}_len"] = np.array([len(r) for r in rays])
            out2[f"s{s}_rays_hg{int(hg)}"] = np.concatenate(rays)
        out2[f"s{s}_rec"] = rec
    qv = queries(rng, origin2, d2, shape2, 30)
    out2["vq"] = qv
    out2["v_interp"] = eik2(qv)
    tts = eik2.solve(srcs2[:2], nsweep=2)
    out2["batch_tt"] = np.stack([t.grid for t in tts])
    print(f"2D done in {time.time() - t0:.1f} s", flush=True)

    np.savez_compressed(os.path.join(HERE, "golden_3d.npz"), **out3)
    np.savez_compressed(os.path.join(HERE, "golden_2d.npz"), **out2)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
