"""Autotuner for the 3D sweep: brick height (``tile_a``) and sources per workgroup (``group``).

The batched 3D solve cuts the grid into bricks of ``tile_a`` z planes x 8 x 8 nodes, one
64-lane wave per brick and source, ``group`` sources per workgroup.  The best pair depends
on the workload:

- tall bricks lose least to the 14-step lane skew in the dense first sweep sets, short
  bricks skip more of the sparse later sets (a brick is skipped only when none of its
  nodes or upwind neighbours changed in the last 8 passes);
- a workgroup's waves meet at every step, so wide groups wait for their slowest source;
- a single source is latency bound and prefers short bricks;
- brick height 0 is the library default (``default_tile``): from 32 sources the height in
  20..28 with the fewest steps down a column, 16 planes from 16 sources, else 8.

``tune3d`` times converged solves of the caller's own workload for a list of candidates
(first the brick heights at a fixed group, then the groups at the best height) and records
the fastest pair for the key (grid shape, number of sources).  ``Eikonal3D.solve``,
``Eikonal3D.solve_tensors`` and the adjoint's forward solve use a recorded pair whenever
``tile_a`` and ``group`` are both left at 0; otherwise the library defaults apply.
Choices can be saved to and loaded from a JSON file.
"""
import json
import time

import numpy as np

_CHOICES = {}

TILES = (8, 16, 32, 48, -1, 0)  # -1: the whole z column; 0: the library default
GROUPS = (1, 2, 4)


def _key(shape, nsrc):
    return "x".join(str(int(n)) for n in shape) + f":{int(nsrc)}"


def lookup3d(shape, nsrc):
    """Recorded (tile_a, group) for a grid of `shape` cells and `nsrc` sources, or (0, 0)."""
    return tuple(_CHOICES.get(_key(shape, nsrc), (0, 0)))


def record3d(shape, nsrc, tile_a, group):
    """Record a choice by hand (e.g. from an earlier tuning run)."""
    _CHOICES[_key(shape, nsrc)] = (int(tile_a), int(group))


def clear():
    _CHOICES.clear()


def save(path):
    with open(path, "w") as f:
        json.dump(_CHOICES, f, indent=1, sort_keys=True)


def load(path):
    with open(path) as f:
        for k, v in json.load(f).items():
            _CHOICES[k] = (int(v[0]), int(v[1]))


def default3d(nsrc):
    """The library's defaults for nsrc sources (fteik3d_solve with tile_a = group = 0):
    brick height 0 (``default_tile``), 2 sources per workgroup from 8 sources."""
    return 0, (2 if nsrc >= 8 else 1)


def default_tile(nsrc, nz_nodes):
    """The brick height the library uses for tile_a = 0 (sweep3d.hip default_tile): from
    32 sources the height t in 20..28 with the fewest steps down a column of nz_nodes node
    planes, ceil(nz_nodes / t) * (t + 14) (the taller on ties); 16 from 16 sources, else 8."""
    if nsrc < 32:
        return 16 if nsrc >= 16 else 8
    best, best_steps = 24, None
    for t in range(20, 29):
        steps = -(-nz_nodes // t) * (t + 14)
        if best_steps is None or steps <= best_steps:
            best, best_steps = t, steps
    return best


def tune3d(eik, sources, *, tiles=TILES, groups=GROUPS, max_sets=100, use_graph=True,
           repeats=1, record=True, margin=0.02):
    """Time converged solves of `sources` on the Eikonal3D `eik` and record the fastest
    (tile_a, group); the library default is kept unless a candidate beats it by more than
    `margin` (a fraction of its time).  Returns {"tile_a", "group", "seconds",
    "timings": {"tile_a:group": s}}.

    Every candidate computes the same traveltimes bit for bit (the sweep order within a
    pass does not change the result); only the time differs.  The solves write into one
    preallocated output, so tuning needs no memory beyond one solve's."""
    import torch

    src = np.asarray(sources, dtype=np.float64).reshape(-1, 3)
    n = len(src)
    nz, nx, ny = eik.shape
    tiles = sorted({min(int(t), nz + 1) if t > 0 else (nz + 1 if t < 0 else 0) for t in tiles})
    groups = [g for g in groups if g <= max(1, n)] or [1]
    out = torch.empty((n, nz + 1, nx + 1, ny + 1), dtype=torch.float64,
                      device=torch.device("cuda", torch.cuda.current_device()))
    timings = {}
    t_start = time.perf_counter()

    def run(tile, group):
        if (tile, group) in timings:
            return timings[(tile, group)]
        best = float("inf")
        for _ in range(max(1, repeats)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            eik.solve_tensors(src, nsweep=None, max_sets=max_sets, tile_a=tile, group=group,
                              use_graph=use_graph, out=out)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        timings[(tile, group)] = best
        return best

    dt, dg = default3d(n)
    g0 = dg if dg in groups else groups[0]
    run(tiles[0], g0)  # warm-up: first launches, graph machinery
    timings.clear()
    best_tile = min(tiles, key=lambda t: run(t, g0))
    best_group = min(groups, key=lambda g: run(best_tile, g))
    if dt in tiles and dg in groups:
        if run(dt, dg) <= timings[(best_tile, best_group)] * (1.0 + margin):
            best_tile, best_group = dt, dg
    del out
    if record:
        record3d(eik.shape, n, best_tile, best_group)
    return {"tile_a": best_tile, "group": best_group,
            "seconds": time.perf_counter() - t_start,
            "timings": {f"{t}:{g}": s for (t, g), s in timings.items()}}
