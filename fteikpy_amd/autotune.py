"""Autotuner for the 3D sweep: brick height (``tile_a``), sources per workgroup (``group``) and
the launch schedule; and, for a batch of models, every 3D model's brick height.

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

- the schedule: "overlapped" (each launch runs the bricks, of any pass, whose inputs are
  final), "diagonal" (one launch per pass and brick anti-diagonal) or "persistent" (one
  cooperative launch per sweep set, a grid-wide barrier between the overlapped list's
  entries); all give the same traveltimes bit for bit.

``tune3d`` times converged solves of the caller's own workload for a list of candidates
(first the brick heights at a fixed group, then the groups at the best height, then the
schedules at the best pair) and records the fastest triple for the key (grid shape, number
of sources).  ``Eikonal3D.solve`` and ``Eikonal3D.solve_tensors`` use a recorded triple
whenever ``tile_a``, ``group`` and ``schedule`` are left at their defaults, the adjoint's
forward solve the recorded height and group; otherwise the library defaults apply.

``tune_batch`` does the same for a batch of models (``solve_batch``): every 3D model gets
its own brick height (the models share every launch, so a height that suits a model alone
need not suit it in the batch).  It times whole-batch rules first (every model at one
height, or the library's rule), then tries other heights for the largest models one at a
time, and records the per-model heights for the batch's signature (each problem's
dimension, grid shape and number of sources); ``solve_batch`` uses them when ``tiles`` is
left unset.  Choices can be saved to and loaded from a JSON file.
"""
import json
import time

import numpy as np

_CHOICES = {}
_BATCH = {}

TILES = (8, 16, 32, 48, -1, 0)  # -1: the whole z column; 0: the library default
GROUPS = (1, 2, 4)
SCHEDULES = ("overlapped", "diagonal", "persistent")  # the first is the library default
BATCH_TILES = (0, 8, 16, 24, 32)  # 0: the library's rule for a batch


def _key(shape, nsrc):
    return "x".join(str(int(n)) for n in shape) + f":{int(nsrc)}"


def lookup3d(shape, nsrc):
    """Recorded (tile_a, group) for a grid of `shape` cells and `nsrc` sources, or (0, 0)."""
    return tuple(_CHOICES.get(_key(shape, nsrc), (0, 0, None))[:2])


def choice3d(shape, nsrc):
    """Recorded (tile_a, group, schedule) for a grid of `shape` cells and `nsrc` sources, or
    (0, 0, None)."""
    return tuple(_CHOICES.get(_key(shape, nsrc), (0, 0, None)))


def record3d(shape, nsrc, tile_a, group, schedule=None):
    """Record a choice by hand (e.g. from an earlier tuning run); schedule None leaves the
    library default."""
    if schedule is not None and schedule not in SCHEDULES:
        raise ValueError(f"schedule must be one of {SCHEDULES}")
    _CHOICES[_key(shape, nsrc)] = (int(tile_a), int(group), schedule)


def batch_signature(problems):
    """The key of a batch: each problem's dimension, grid shape and number of sources."""
    sig = []
    for eik, sources in problems:
        nd = len(eik.shape)
        ns = len(np.asarray(sources, dtype=np.float64).reshape(-1, nd))
        sig.append(f"{nd}d" + "x".join(str(int(n)) for n in eik.shape) + f":{ns}")
    return ";".join(sig)


def lookup_batch(problems):
    """Recorded per-problem brick heights for this batch (0 for 2D problems), or None."""
    v = _BATCH.get(batch_signature(problems))
    return list(v) if v is not None else None


def record_batch(problems, tiles):
    tiles = [int(t) for t in tiles]
    if len(tiles) != len(problems):
        raise ValueError("one brick height per problem")
    _BATCH[batch_signature(problems)] = tiles


def clear():
    _CHOICES.clear()
    _BATCH.clear()


def save(path):
    with open(path, "w") as f:
        json.dump({"single": {k: list(v) for k, v in _CHOICES.items()}, "batch": _BATCH}, f,
                  indent=1, sort_keys=True)


def load(path):
    with open(path) as f:
        data = json.load(f)
    single = data.get("single", {}) if "single" in data or "batch" in data else data
    for k, v in single.items():
        sched = v[2] if len(v) > 2 else None
        _CHOICES[k] = (int(v[0]), int(v[1]), sched)
    for k, v in data.get("batch", {}).items():
        _BATCH[k] = [int(t) for t in v]


def default3d(nsrc):
    """The library's defaults for nsrc sources (fteik3d_solve with tile_a = group = 0):
    brick height 0 (``default_tile``); group 0 from 32 sources (4 sources per workgroup in
    the first 2 sweep sets, 2 in the later ones), else 2 sources per workgroup from 8
    sources, 1 below."""
    if nsrc >= 32:
        return 0, 0
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


def tune3d(eik, sources, *, tiles=TILES, groups=GROUPS, schedules=SCHEDULES, max_sets=100,
           use_graph=True, repeats=1, record=True, margin=0.02):
    """Time converged solves of `sources` on the Eikonal3D `eik` and record the fastest
    (tile_a, group, schedule); the library default is kept unless a candidate beats it by
    more than `margin` (a fraction of its time).  Returns {"tile_a", "group", "schedule",
    "seconds", "timings": {"tile_a:group:schedule": s}}.

    Every candidate computes the same traveltimes bit for bit (the sweep order within a
    pass does not change the result); only the time differs.  The solves write into one
    preallocated output, so tuning needs no memory beyond one solve's."""
    import torch

    src = np.asarray(sources, dtype=np.float64).reshape(-1, 3)
    n = len(src)
    nz, nx, ny = eik.shape
    tiles = sorted({min(int(t), nz + 1) if t > 0 else (nz + 1 if t < 0 else 0) for t in tiles})
    groups = [g for g in groups if g <= max(1, n)] or [1]
    dt, dg = default3d(n)
    if dg not in groups:
        groups = [dg] + groups  # the library's default group is always a candidate
    out = torch.empty((n, nz + 1, nx + 1, ny + 1), dtype=torch.float64,
                      device=torch.device("cuda", torch.cuda.current_device()))
    schedules = [s for s in schedules if s in SCHEDULES] or [SCHEDULES[0]]
    s0 = schedules[0]
    timings = {}
    t_start = time.perf_counter()

    def run(tile, group, sched=s0):
        if (tile, group, sched) in timings:
            return timings[(tile, group, sched)]
        best = float("inf")
        for _ in range(max(1, repeats)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            eik.solve_tensors(src, nsweep=None, max_sets=max_sets, tile_a=tile, group=group,
                              schedule=sched, use_graph=use_graph, out=out)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        timings[(tile, group, sched)] = best
        return best

    g0 = dg
    run(tiles[0], g0)  # warm-up: first launches, graph machinery
    timings.clear()
    best_tile = min(tiles, key=lambda t: run(t, g0))
    best_group = min(groups, key=lambda g: run(best_tile, g))
    best_sched = min(schedules, key=lambda sc: run(best_tile, best_group, sc))
    ds = SCHEDULES[0]
    if dt in tiles and dg in groups and ds in schedules:
        if run(dt, dg, ds) <= timings[(best_tile, best_group, best_sched)] * (1.0 + margin):
            best_tile, best_group, best_sched = dt, dg, ds
    del out
    if record:
        record3d(eik.shape, n, best_tile, best_group, best_sched)
    return {"tile_a": best_tile, "group": best_group, "schedule": best_sched,
            "seconds": time.perf_counter() - t_start,
            "timings": {f"{t}:{g}:{sc}": s for (t, g, sc), s in timings.items()}}


def tune_batch(problems, *, heights=BATCH_TILES, refine=4, group=0, max_sets=100,
               use_graph=True, record=True, margin=0.02):
    """Time converged batch solves (``solve_batch_tensors``) of `problems` and record the
    fastest brick height of every 3D model; the library's rule is kept unless a candidate
    beats it by more than `margin`.  Candidates: every 3D model at each height of `heights`
    (0 = the library's rule for a batch: the height for the batch's total 3D source count),
    then, from the best of those, each of the `refine` largest 3D models (grid nodes x
    sources) at every other height, one model at a time, kept when faster.  Every candidate
    gives the same traveltimes bit for bit.  Returns {"tiles", "seconds", "timings"}."""
    import torch

    from ._batch import solve_batch_tensors

    is3 = [len(eik.shape) == 3 for eik, _ in problems]
    idx3 = [i for i, b in enumerate(is3) if b]
    t_start = time.perf_counter()
    timings = {}

    def run(tiles):
        key = tuple(tiles)
        if key in timings:
            return timings[key]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = solve_batch_tensors(problems, nsweep=None, max_sets=max_sets, group=group,
                                  use_graph=use_graph, tiles=list(tiles))
        torch.cuda.synchronize()
        timings[key] = time.perf_counter() - t0
        del res
        return timings[key]

    def uniform(h):
        return tuple(h if b else 0 for b in is3)

    base = uniform(0)
    if not idx3:
        best = base
    else:
        run(base)  # warm-up
        timings.clear()
        best = min((uniform(h) for h in heights), key=run)

        def size(i):
            eik, src = problems[i]
            ns = len(np.asarray(src, dtype=np.float64).reshape(-1, 3))
            return np.prod([n + 1 for n in eik.shape]) * ns

        for i in sorted(idx3, key=size, reverse=True)[:max(0, refine)]:
            for h in heights:
                cand = list(best)
                cand[i] = h
                cand = tuple(cand)
                if run(cand) < run(best):
                    best = cand
        if run(base) <= run(best) * (1.0 + margin):
            best = base
    if record:
        record_batch(problems, best)
    return {"tiles": list(best), "seconds": time.perf_counter() - t_start,
            "timings": {",".join(map(str, k)): v for k, v in timings.items()}}
