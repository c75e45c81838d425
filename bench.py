"""Benchmark: converged FP64 3D source-grids per second, batched vs plain, on one GPU; and
seconds per forward+adjoint iteration of the gradient workload.

Workload (default): 256 sources on a 160 x 320 x 320 FP64 model, 12.5 m spacing, linear
gradient 2000 + 0.8 z m/s with 20 spherical inclusions (+-15%); sources on a 16 x 16
surface grid at 6.25 m depth.  A step is one batched solve of all sources swept to
convergence (sets repeat until one changes no node), sweep sets replayed from a hipGraph.

"plain" is the per-source solve looped over a subset of the sources.  The gradient
workload (fwd_adjoint in the output) is described in fwd_adjoint().  The CPU baseline is
the oracle (C restatement of the reference's numba kernels, OpenMP over sources like the
reference's prange) timed on this host on a bounded subset of the same sources.
Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# LDS: 128 bytes per clock per CU, 256 CUs, 2.4 GHz peak engine clock
LDS_PEAK_GBS = 128 * 256 * 2.4
BYTES_PER_NSU = 24  # FP64 tt read + tt write + slowness read per node update

# HBM bytes per launch of the dominant kernel, sweep3d_diag_kernel (<false, false, 4> in the
# first two sets, <false, false, 2> after them), on the default workload (160 x 320 x 320,
# 256 sources, the library's default brick height, 27 planes here, and group, the
# overlapped schedule): rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of this
# script, averaged over the 19812 launches of one converged solve and scaled by a 1 GiB
# device copy (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0).  Commands and figures:
# profiles/sweep3d_traffic__inclusions_160x320x320_256src.md
PMC_TRAFFIC = {"grid": (160, 320, 320), "sources": 256, "tile_a": 0, "group": 0,
               "fetch_bytes": 208725.2 * 2.0 * 1024, "write_bytes": 11468.6 * 1.0 * 1024,
               # SQ_INSTS_LDS per launch x 64 lanes x 5.4 B, the mean bytes per lane of the
               # step loop's LDS instructions (8 ds_read2_b64, 2 ds_read_b64, 6 ds_write_b64, 18
               # ds_read_u8, 6 ds_write_b8 per two unrolled steps)
               "lds_bytes": 2.917366e6 * 64 * 5.4}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nsrc", type=int, default=256)
    ap.add_argument("--plain", type=int, default=4, help="sources timed in the plain loop")
    ap.add_argument("--plain-tile", type=int, default=0)
    ap.add_argument("--tile", type=int, default=0, help="batched brick height (0 = default)")
    ap.add_argument("--group", type=int, default=0,
                    help="batched sources per workgroup (0 = the autotuner's choice or default)")
    ap.add_argument("--shape", type=str, default="160,320,320")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the autotuner (brick height and group from the library defaults)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sources", type=int, default=0, help="0 = one per host core (max 16)")
    ap.add_argument("--cpu-sets", type=int, default=4, help="sweep sets timed on the CPU")
    ap.add_argument("--adj-steps", type=int, default=3,
                    help="timed forward+adjoint iterations of the gradient workload (0 = skip)")
    ap.add_argument("--adj-shape", type=str, default="128,256,256")
    ap.add_argument("--no-analytic", action="store_true",
                    help="skip the error report against closed-form traveltimes")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the uneven multi-model batch (48 models, 16 2D + 32 3D)")
    ap.add_argument("--profile-c2", action="store_true",
                    help="only run the single-source 200^3 solve (a warm-up and 3 converged "
                         "solves) and print its times, for profiling")
    ap.add_argument("--profile-adjoint", action="store_true",
                    help="only run the gradient workload (one warm-up and one timed "
                         "forward+adjoint iteration), for profiling")
    ap.add_argument("--profile-batch", action="store_true",
                    help="only run the uneven multi-model batch (warm-up and one timed call of "
                         "all 48 models), for profiling")
    ap.add_argument("--profile-plain", action="store_true",
                    help="only run the per-source loop of the batched workload (a warm-up and "
                         "--plain converged single-source solves), for profiling")
    ap.add_argument("--profile-post", action="store_true",
                    help="only run the batched workload's solve with gradient signs and its "
                         "per-source gradient grids, receivers and rays, for profiling")
    ap.add_argument("--profile-2d", action="store_true",
                    help="only run the 2D workloads (the plain 2D case at nsweep 2, the 2D "
                         "models of the uneven batch) and print their times, for profiling")
    ap.add_argument("--profile-schedules", action="store_true",
                    help="only run converged solves of the single-source 200^3 workload and the "
                         "batched workload (brick height and group at the library defaults) on "
                         "the overlapped and the persistent schedule, a warm-up and one timed "
                         "solve each, and print their times, for profiling")
    ap.add_argument("--copy-calibration", type=float, default=0.0, metavar="GIB",
                    help="only copy GIB GiB once on the device (torch clone) and exit: the "
                         "known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--row-calibration", action="store_true",
                    help="only copy 72-byte rows (9 doubles at element offset 7 of every "
                         "321-double row of a 161 x 321 x 321 grid, the sweep kernel's staging "
                         "reads) 8 times and exit: the row-granular reference for FETCH_SIZE")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.row_calibration:
        row_calibration()
        return

    if args.profile_2d:
        profile_2d()
        return
    if args.profile_schedules:
        profile_schedules(args)
        return
    if args.profile_batch or args.profile_plain or args.profile_post:
        profile_sections(args)
        return
    if args.profile_c2 or args.profile_adjoint:
        import torch

        import fteikpy_amd
        from fteikpy_amd import synthetic

        dev = torch.device("cuda", 0)
        with torch.cuda.stream(torch.cuda.Stream(device=dev)):
            if args.profile_c2:
                out = single_source_c2(torch, fteikpy_amd, synthetic, dev, with_cpu=False)
            else:
                args.adj_steps, args.no_tune = 1, True
                out = fwd_adjoint(args, torch, fteikpy_amd, synthetic, dev)
        print(json.dumps(out), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); sources are independent, "
              "see DESIGN.md for how the path shards", file=sys.stderr)
        sys.exit(2)

    import torch

    import fteikpy_amd
    from fteikpy_amd import synthetic

    shape = tuple(int(s) for s in args.shape.split(","))
    nside = int(round(args.nsrc ** 0.5))
    v, d, src, rec = synthetic.model_3d_inclusions(shape=shape, nsrc_side=nside)
    src = src[: args.nsrc]
    nsrc = len(src)
    nz, nx, ny = shape
    dev = torch.device("cuda", 0)
    eik = fteikpy_amd.Eikonal3D(v, d)
    stream = torch.cuda.Stream(device=dev)
    use_graph = not args.no_graph

    tune = None
    with torch.cuda.stream(stream):
        if not args.no_tune and not args.tile and not args.group:
            # brick height, group and schedule for this workload, timed on it before the
            # warm-up; the solves below leave all three at their defaults and so use them
            tune = fteikpy_amd.autotune.tune3d(eik, src, tiles=(16, 32, 0),
                                               use_graph=use_graph)
        if tune:
            log(f"autotune: tile_a {tune['tile_a']} group {tune['group']} schedule "
                f"{tune['schedule']} ({tune['seconds']:.1f} s)")
        out = torch.empty((nsrc, nz + 1, nx + 1, ny + 1), dtype=torch.float64, device=dev)
        for _ in range(args.warmup):
            eik.solve_tensors(src, nsweep=None, tile_a=args.tile, group=args.group,
                              use_graph=use_graph, out=out)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        sets_total = 0
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            tt, _, vz, sets = eik.solve_tensors(src, nsweep=None, tile_a=args.tile,
                                                group=args.group, use_graph=use_graph, out=out)
            sets_total += int(sets.sum())
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gpu_s = ev0.elapsed_time(ev1) / 1e3
        log(f"batched: {args.steps} converged solves in {wall:.2f} s")
        batched_tt0 = tt[0].cpu().numpy()
        k_sets = sets.copy()

        # plain: one source at a time
        nplain = min(args.plain, nsrc)
        plain_s, plain_equal = None, None
        if nplain > 0:
            eik.solve_tensors(src[:1], nsweep=None, tile_a=args.plain_tile)  # warm
            torch.cuda.synchronize()
            tp = time.perf_counter()
            plain_tt0 = None
            for i in range(nplain):
                ttp, _, _, _ = eik.solve_tensors(src[i:i + 1], nsweep=None,
                                                 tile_a=args.plain_tile)
                if i == 0:
                    plain_tt0 = ttp[0]
            torch.cuda.synchronize()
            plain_s = time.perf_counter() - tp
            plain_equal = bool(np.array_equal(plain_tt0.cpu().numpy(), batched_tt0))
            log(f"plain: {nplain} solves in {plain_s:.2f} s")

    del out, tt
    torch.cuda.empty_cache()
    with torch.cuda.stream(stream):
        post = None if args.no_analytic else headline_post(torch, fteikpy_amd, eik, src, rec, d)
    torch.cuda.empty_cache()
    # every section runs on the one stream above: HIP maps a process's streams onto at most
    # GPU_MAX_HW_QUEUES hardware queues (4 by default), and the multi-model batch overlaps
    # its 2D and 3D parts only while its side stream has a queue of its own
    with torch.cuda.stream(stream):
        adj = (fwd_adjoint(args, torch, fteikpy_amd, synthetic, dev) if args.adj_steps > 0
               else None)
        ana = None if args.no_analytic else analytic_errors(torch, fteikpy_amd, dev)
        c2 = None if args.no_analytic else single_source_c2(torch, fteikpy_amd, synthetic, dev,
                                                            with_cpu=not args.no_cpu)
        ub = None if args.no_batch else uneven_batch(torch, fteikpy_amd, synthetic, dev,
                                                     tune=not args.no_tune)
        p2d = None if args.no_analytic else plain_2d(torch, fteikpy_amd, synthetic, dev,
                                                     with_cpu=not args.no_cpu)

    nodes = nz * nx * ny
    alg_bytes = sets_total * 8 * nodes * BYTES_PER_NSU
    achieved = alg_bytes / gpu_s / 1e9
    value = nsrc * args.steps / wall
    # launches of the sweep kernel in the timed region: every set of a batch makes the
    # same launches (the library's count for this grid, brick height and the default
    # overlapped schedule), until its last source converged
    tile = args.tile or (tune["tile_a"] if tune else 0)
    group = args.group or (tune["group"] if tune else 0)  # 0: the library's default
    sched = tune["schedule"] if tune else "overlapped"
    per_set = fteikpy_amd._lib.lib().fteik3d_sweep_launches(
        nz, nx, ny, nsrc, tile, {"overlapped": 2, "diagonal": 1}[sched])
    launches = args.steps * int(k_sets.max()) * per_set
    traffic = None
    if (tuple(shape) == PMC_TRAFFIC["grid"] and nsrc == PMC_TRAFFIC["sources"]
            and tile == PMC_TRAFFIC["tile_a"] and group == PMC_TRAFFIC["group"]):
        traffic = PMC_TRAFFIC["fetch_bytes"] + PMC_TRAFFIC["write_bytes"]

    cpu = None
    if not args.no_cpu:
        import oracle

        ncores, host = host_cpu()
        ncs = args.cpu_sources or ncores
        nsets = args.cpu_sets
        tc = time.perf_counter()
        oracle.solve3d(1.0 / v, *d, src[:ncs], nsweep=nsets, nthreads=ncores)
        tcpu = time.perf_counter() - tc
        log(f"cpu baseline: {tcpu:.1f} s on {ncores} threads ({host['model']})")
        # the CPU path evaluates every node in every set, so a converged grid costs its set
        # count (measured above for the same sources) times the per-set time
        mean_sets = float(k_sets[:ncs].mean())
        cpu = {"value": ncs * nsets / tcpu / mean_sets, "unit": "source-grids/s",
               "cores": ncores, "kind": "port", "host": host,
               "sample": f"{ncs} of the {nsrc} sources x {nsets} sweep sets in {tcpu:.1f} s on "
                         f"{ncores} threads, scaled to the {mean_sets:.1f} sets those sources "
                         "need to converge; oracle C restatement of the reference's numba "
                         "path (its k,j,i loop nesting), OpenMP over sources like the "
                         "reference's prange"}

    torch.cuda.empty_cache()
    copy_gbs = stream_copy(torch, dev, stream)
    lds_b = PMC_TRAFFIC["lds_bytes"] if traffic else None
    us_launch = gpu_s / launches * 1e6
    res = {
        "metric": "converged FP64 source-grids/sec 3D, batched vs plain GPU; fwd+adjoint s/iter",
        "value": value,
        "unit": "source-grids/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "batched many-source 3D eikonal, linear gradient + 20 spherical "
                               "inclusions, converged",
                   "grid": list(shape), "spacing_m": d[0], "sources": nsrc,
                   "sets_per_source": [int(k_sets.min()), int(k_sets.max())],
                   "hipgraph": use_graph,
                   "tile_a": args.tile or (tune["tile_a"] if tune else 0),
                   "group": args.group or (tune["group"] if tune else 0),
                   "schedule": sched},
        "autotune": tune,
        "plain_value": nplain / plain_s if plain_s else None,
        "plain_sources_timed": nplain,
        "batched_over_plain": value / (nplain / plain_s) if plain_s else None,
        "batched_equals_plain_bitwise": plain_equal,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": alg_bytes / launches,
                     "launches": launches, "us_per_launch": us_launch,
                     "traffic_GBps": traffic / us_launch / 1e3 if traffic else None,
                     "frac_traffic": traffic / us_launch / 1e3 / HBM_PEAK_GBS if traffic else None,
                     "lds_bytes_per_launch_est": lds_b,
                     "lds_GBps_est": lds_b / us_launch / 1e3 if lds_b else None,
                     "lds_peak_GBps": LDS_PEAK_GBS,
                     "stream_copy_GBps": copy_gbs, "achieved_over_stream_copy": achieved / copy_gbs,
                     "note": "achieved (frac) = WORK-EQUIVALENT bandwidth: algorithmic bytes "
                             "per launch (24 B per node update x 8 passes x nodes x sets "
                             "executed, over the launches of the timed solves, counting the "
                             "node updates the kernel proves unnecessary and skips) / mean "
                             "launch time (HIP events on the kernels' stream); traffic = "
                             "measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                             "calibrated on copies of known size; null for other "
                             "configurations) and frac_traffic = traffic / mean launch time / "
                             "peak, the bandwidth actually used; lds_bytes_per_launch_est = "
                             "LDS instructions per launch (rocprofv3 SQ_INSTS_LDS) x 64 lanes x "
                             "the mean bytes per lane of the kernel's LDS instruction mix "
                             "(an upper bound: inactive lanes counted); stream_copy_GBps = a "
                             "1 GiB device copy (read + write bytes) timed here"},
        "cpu_baseline": cpu,
        "fwd_adjoint": adj,
        "analytic": ana,
        "single_source_200cubed": c2,
        "headline_post": post,
        "uneven_batch": ub,
        "plain_2d_101x301": p2d,
    }
    print(json.dumps(res), flush=True)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cpu():
    """(threads used, description) of the host CPU: every logical CPU this process may run
    on (its affinity mask), capped by OMP_NUM_THREADS when the environment sets a share for
    this process, and the CPU model from /proc/cpuinfo."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS")
    used = min(avail, int(share)) if share and share.isdigit() and int(share) > 0 else avail
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return used, {"model": model, "logical_cpus_available": avail,
                  "omp_num_threads": share, "threads_used": used}


def copy_calibration(gib):
    """One device copy of `gib` GiB (read once, written once): the reference kernel for the
    HBM byte counters."""
    import torch

    n = int(gib * (1 << 30)) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    y = x.clone()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "seconds": dt,
                      "checksum": float(y[::4096].sum())}), flush=True)


def row_calibration(reps=8, shape=(161, 321, 321), first=7):
    """Copies of 72-byte rows: every (z, x) row of a traveltime-sized grid contributes
    elements first..first+8 (the sweep kernel stages 9 x 9 node blocks whose rows start one
    element before a multiple of 8), each copy one kernel reading z*x*72 bytes in 72-byte
    pieces and writing them contiguously."""
    import torch

    g = torch.ones(shape, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    out = None
    for _ in range(reps):
        out = g[:, :, first:first + 9].clone()
    torch.cuda.synchronize()
    rows = shape[0] * shape[1]
    print(json.dumps({"row_copy_dispatches": reps, "bytes_read_per_dispatch": rows * 72,
                      "bytes_written_per_dispatch": rows * 72,
                      "checksum": float(out.sum())}), flush=True)


def headline_post(torch, fa, eik, src, rec, d, nrays=64, chunk=32):
    """The rest of the batched configuration: the converged solve again with gradient
    signs, then per source its gradient grids, traveltimes at the 4096 receivers and
    `nrays` ray paths (to the first receivers), one source after another (wall time)."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    tt, sgn, vz, sets = eik.solve_tensors(src, nsweep=None, return_gradient=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    npts, first = 0, None
    for i in range(len(src)):
        g = fa.TraveltimeGrid3D(tt[i], d, np.zeros(3), src[i], None, float(vz[i]), sign=sgn[i])
        g(rec)
        rays = g.raytrace(rec[:nrays])
        npts += sum(len(r) for r in rays)
        if i == 0:
            first = rays
        del g
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    # the same with the rays of `chunk` sources traced in one launch (raytrace_batch); the
    # receivers stay one call per source
    npts_b, equal = 0, None
    for c0 in range(0, len(src), chunk):
        grids = [fa.TraveltimeGrid3D(tt[i], d, np.zeros(3), src[i], None, float(vz[i]),
                                     sign=sgn[i]) for i in range(c0, min(len(src), c0 + chunk))]
        for g in grids:
            g(rec)
        rays = fa.raytrace_batch(grids, [rec[:nrays]] * len(grids))
        npts_b += sum(len(r) for rr in rays for r in rr)
        if c0 == 0:
            equal = all(np.array_equal(a, b) for a, b in zip(rays[0], first))
        del grids, rays
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    del tt, sgn
    log(f"headline post: solve with signs {t1 - t0:.2f} s, receivers and rays {t2 - t1:.2f} s "
        f"(rays batched by {chunk}: {t3 - t2:.2f} s)")
    return {"solve_with_signs_s": t1 - t0, "receivers_and_rays_s": t2 - t1,
            "receivers_and_rays_batched_s": t3 - t2, "rays_batch_sources": chunk,
            "batched_rays_equal_bitwise": bool(equal) and npts_b == npts,
            "receivers_per_source": len(rec), "rays_per_source": nrays,
            "ray_points": npts,
            "note": "per source: gradient grids on the device (from the signs), "
                    "traveltimes at its receivers, rays to the first receivers (returned "
                    "to the host as the reference's lists of arrays); batched: the rays of "
                    f"{chunk} sources per raytrace_batch call (one launch), the same rays"}


def single_source_c2(torch, fa, synthetic, dev, reps=3, with_cpu=True):
    """The single-source configuration: 200^3 layered model with a smooth perturbation,
    5 m, one off-grid source, converged solve (tt stays on the device).  Best of `reps`
    after a warm-up, with its algorithmic rate (24 B per node update).  Beside it the
    oracle (the reference's numba path restated in C, its loop nesting) on one host thread
    (one source: the reference's prange over sources has nothing to spread): one sweep set
    timed, scaled to the set count of the converged solve (the CPU path evaluates every
    node in every set)."""
    v, d, src, rec = synthetic.model_3d_layers()
    eik = fa.Eikonal3D(v, d)
    eik.solve_tensors(src, nsweep=None)
    torch.cuda.synchronize()
    best, sets = None, 0
    for _ in range(reps):
        t0 = time.perf_counter()
        _, _, _, k = eik.solve_tensors(src, nsweep=None)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        best = t if best is None else min(best, t)
        sets = int(k[0])
    nodes = int(np.prod(np.array(v.shape)))
    alg = sets * 8 * nodes * BYTES_PER_NSU
    # the whole configuration: converged solve with gradient signs, the gradient grids and
    # traveltimes at the 10,000 receivers
    full = None
    for _ in range(reps):
        t0 = time.perf_counter()
        g = eik.solve(src, nsweep=None, return_gradient=True)
        g._gradient_tensor()  # the gradient grids, on the device
        g(rec)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        full = t if full is None else min(full, t)
    cpu = None
    if with_cpu:
        import oracle

        tc = time.perf_counter()
        oracle.solve3d(1.0 / v, *d, src, nsweep=1, nthreads=1)
        t1 = time.perf_counter() - tc
        _, host = host_cpu()
        cpu = {"value": 1.0 / (t1 * sets), "unit": "source-grids/s", "cores": 1,
               "kind": "port", "host_model": host["model"],
               "sample": f"one sweep set of the source in {t1:.1f} s on one thread, scaled to "
                         f"the {sets} sets of the converged solve"}
        log(f"single source: cpu {t1:.1f} s per set")
    return {"grid": list(v.shape), "spacing_m": d[0], "sets": sets, "seconds": best,
            "cpu_baseline": cpu,
            "source_grids_per_s": 1.0 / best, "achieved_GBps": alg / best / 1e9,
            "seconds_with_gradient_and_10000_receivers": full,
            "note": "wall time of one converged solve (seeding, sets, convergence checks), "
                    "brick height and group at the library defaults for one source; then "
                    "of the same solve with gradient signs, the gradient grids and the "
                    "traveltimes at 10,000 receivers"}


def profile_sections(args):
    """One bench section alone, on one stream, for a profiler: the uneven batch (a warm-up
    and one timed call), the per-source loop of the batched workload (a warm-up and
    args.plain solves), or the batched workload's solve with signs and its per-source
    gradients, receivers and rays.  Prints the section's JSON."""
    import torch

    import fteikpy_amd as fa
    from fteikpy_amd import synthetic

    dev = torch.device("cuda", 0)
    with torch.cuda.stream(torch.cuda.Stream(device=dev)):
        if args.profile_batch:
            out = uneven_batch(torch, fa, synthetic, dev, reps=1)
        else:
            shape = tuple(int(x) for x in args.shape.split(","))
            v, d, src, rec = synthetic.model_3d_inclusions(shape=shape)
            eik = fa.Eikonal3D(v, d)
            if args.profile_post:
                out = headline_post(torch, fa, eik, src, rec, d)
            else:
                n = max(1, args.plain)
                eik.solve_tensors(src[:1], nsweep=None)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                sets = []
                for i in range(n):
                    _, _, _, k = eik.solve_tensors(src[i:i + 1], nsweep=None)
                    sets.append(int(k[0]))
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                out = {"plain_sources": n, "seconds": dt, "source_grids_per_s": n / dt,
                       "sets": sets}
    print(json.dumps(out), flush=True)


def profile_schedules(args):
    """Overlapped (graph-replayed launches) against persistent (one cooperative launch per
    sweep set) on the single-source 200^3 workload and the batched workload: a warm-up and
    one timed converged solve each, on one stream; prints the times and whether the
    persistent solves equal the overlapped ones bit for bit."""
    import torch

    import fteikpy_amd as fa
    from fteikpy_amd import synthetic

    dev = torch.device("cuda", 0)
    shape = tuple(int(x) for x in args.shape.split(","))
    cases = [("single_source_200cubed", synthetic.model_3d_layers()),
             ("batched", synthetic.model_3d_inclusions(shape=shape))]
    out = {}
    with torch.cuda.stream(torch.cuda.Stream(device=dev)):
        for name, (v, d, src, _) in cases:
            eik = fa.Eikonal3D(v, d)
            res = {}
            ref = None
            for sched in ("overlapped", "persistent"):
                def solve():
                    return eik.solve_tensors(src, nsweep=None, tile_a=0, group=0, schedule=sched,
                                             use_graph=sched == "overlapped")
                solve()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                tt, _, _, sets = solve()
                torch.cuda.synchronize()
                res[sched] = time.perf_counter() - t0
                log(f"{name} {sched}: {res[sched]:.4f} s")
                if ref is None:
                    ref = (tt, list(sets))
                else:
                    res["persistent_equals_overlapped_bitwise"] = bool(
                        torch.equal(tt, ref[0]) and list(sets) == ref[1])
                    del tt
            del ref
            out[name] = res
    print(json.dumps(out), flush=True)


def profile_2d(reps=2):
    """The 2D workloads alone: the plain 2D case (one source, nsweep 2, gradient signs, 200
    receivers) and the 16 2D models of the uneven batch converged, `reps` times each after
    a warm-up, on one stream."""
    import torch

    import fteikpy_amd as fa
    from fteikpy_amd import synthetic

    dev = torch.device("cuda", 0)
    v, d, src, rec = synthetic.model_2d_gaussians()
    eik = fa.Eikonal2D(v, d)
    p2 = [(fa.Eikonal2D(v2, d2), s2) for v2, d2, s2 in synthetic.uneven_batch(n3=0)]
    out = {}
    with torch.cuda.stream(torch.cuda.Stream(device=dev)):
        for name, fn in (("plain_2d_nsweep2", lambda: eik.solve(src, 2, True)(rec)),
                         ("uneven_batch_2d", lambda: fa.solve_batch_tensors(p2, nsweep=None))):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                fn()
            torch.cuda.synchronize()
            out[name + "_s"] = (time.perf_counter() - t0) / reps
            log(f"{name}: {out[name + '_s']:.4f} s")
    print(json.dumps(out), flush=True)


def plain_2d(torch, fa, synthetic, dev, with_cpu=True, reps=3):
    """The plain 2D case on its one source (101 x 301 cells of 10 m, 1500 + 0.6 z m/s with
    four Gaussian anomalies, source at (12.3, 1013.7) m): solve with gradient signs at
    nsweep=2 and converged, then traveltimes at 200 receivers; GPU wall time (best of
    `reps`) beside the oracle's single-thread C restatement of the reference's numba path
    on the same solve (the 2D chain is serial per source, so one source is latency bound on
    either side); then 512 sources converged, beside the oracle on the host's threads."""
    v, d, src, rec = synthetic.model_2d_gaussians()
    eik = fa.Eikonal2D(v, d)
    out = {}
    for label, ns in (("nsweep2", 2), ("converged", None)):
        ts = []
        for _ in range(reps + 1):
            t0 = time.perf_counter()
            g = eik.solve(src, nsweep=ns, return_gradient=True)
            g(rec)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        entry = {"gpu_s": min(ts[1:])}
        if with_cpu:
            import oracle

            tc = time.perf_counter()
            oracle.solve2d(1.0 / v, *d, src[None, :], nsweep=ns or 0, grad=True, nthreads=1)
            entry["cpu_oracle_1_thread_s"] = time.perf_counter() - tc
        out[label] = entry
    # many sources: 512 uniform in the model, converged
    srcs = np.random.default_rng(3).random((512, 2)) * np.array(v.shape) * d
    ts = []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        eik.solve(srcs, nsweep=None)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    entry = {"gpu_s": min(ts[1:])}
    if with_cpu:
        import oracle

        ncores, _ = host_cpu()
        tc = time.perf_counter()
        oracle.solve2d(1.0 / v, *d, srcs, nsweep=0, nthreads=ncores)
        entry["cpu_oracle_s"] = time.perf_counter() - tc
        entry["cpu_threads"] = ncores
    out["converged_512_sources"] = entry
    return out


def uneven_batch(torch, fa, synthetic, dev, reps=3, tune=False):
    """The uneven batch of independent models (16 2D models 64 to 1024 per side and 32 3D
    models 48^3 to 192 x 256 x 256, 1 to 32 sources each, smooth random velocities),
    converged in one fteik_solve_batch call; beside it the 3D and the 2D models alone and
    the 3D models solved one call each (best of `reps` after a warm-up, wall time).  With
    `tune`, the batch autotuner picks every 3D model's brick height first (recorded for the
    whole batch and for its 3D part) and the batch calls use them."""
    models = synthetic.uneven_batch()
    probs = [((fa.Eikonal3D if v.ndim == 3 else fa.Eikonal2D)(v, d), s) for v, d, s in models]
    p3 = [p for p in probs if isinstance(p[0], fa.Eikonal3D)]
    p2 = [p for p in probs if not isinstance(p[0], fa.Eikonal3D)]
    tuned = None
    if tune:
        tuned = fa.autotune.tune_batch(probs)
        fa.autotune.record_batch(p3, [t for (e, _), t in zip(probs, tuned["tiles"])
                                      if isinstance(e, fa.Eikonal3D)])
        log(f"batch autotune: heights {tuned['tiles']} ({tuned['seconds']:.1f} s)")

    slowest = {}

    def best(fn, n, key):
        r = fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(n):
            del r
            t0 = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        slowest[key] = max(ts)
        return min(ts), r

    # on the caller's current stream
    t_all, res = best(lambda: fa.solve_batch_tensors(probs, nsweep=None), reps, "all")
    t3, _ = best(lambda: fa.solve_batch_tensors(p3, nsweep=None), reps, "3d")
    t2, _ = best(lambda: fa.solve_batch_tensors(p2, nsweep=None), reps, "2d")
    tloop, loop = best(lambda: [e.solve_tensors(s, nsweep=None) for e, s in p3], 1, "loop")
    off = len(p2)
    equal = all(torch.equal(res[off + i][0], loop[i][0]) for i in range(len(p3)))
    sets = [int(r[3].max()) for r in res]
    del res, loop
    torch.cuda.empty_cache()
    log(f"uneven batch: all {t_all:.2f} s, 3D {t3:.2f} s, 2D {t2:.2f} s, "
        f"3D one call each {tloop:.2f} s")
    return {"models_2d": len(p2), "models_3d": len(p3),
            "sources_2d": sum(len(s) for _, s in p2), "sources_3d": sum(len(s) for _, s in p3),
            "seconds_all": t_all, "seconds_3d_only": t3, "seconds_2d_only": t2,
            "seconds_3d_one_call_each": tloop, "batched_3d_over_one_call_each": tloop / t3,
            "batched_equals_one_call_each_bitwise": equal,
            "max_sets_2d": max(sets[:off]), "max_sets_3d": max(sets[off:]),
            "slowest_repetition_s": slowest, "autotune": tuned,
            "note": f"converged (nsweep=None); wall time, best of {reps} repetitions after a "
                    "warm-up (the 3D one-call-each loop: 1); the 2D models run on their own "
                    "stream beside the 3D sweeps"}


def stream_copy(torch, dev, stream, gib=1, reps=10):
    """Device-to-device copy bandwidth (STREAM copy): bytes read + written per second over
    `reps` copies of a `gib` GiB FP64 buffer, timed with HIP events on the copy stream."""
    n = gib * (1 << 27)
    a = torch.empty(n, dtype=torch.float64, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    with torch.cuda.stream(stream):
        for _ in range(2):
            b.copy_(a)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            b.copy_(a)
        e1.record(stream)
    torch.cuda.synchronize()
    secs = e0.elapsed_time(e1) / 1e3
    del a, b
    return 2.0 * n * 8 * reps / secs / 1e9


def analytic_errors(torch, fa, dev):
    """Error of converged solves against closed forms on the 200^3, 5 m grid of the
    single-source configuration (source at (502.3, 497.1, 503.9) m): a homogeneous medium
    (3000 m/s, t = |x - x_s| / v) and a constant vertical gradient v = v0 + g z (1800 m/s
    at the top, g = 2.7 1/s: 1800..4500 m/s over the model; cell velocities at cell
    centres; t = arccosh(1 + g^2 R^2 / (2 v(z_s) v(z))) / g).  Errors over the nodes more
    than 10 cells from the source; FTeik is first order, so they shrink with the spacing."""
    n, h = 200, 5.0
    src = np.array([502.3, 497.1, 503.9])
    ax = torch.arange(n + 1, dtype=torch.float64, device=dev) * h
    z = ax[:, None, None]
    r2 = (z - src[0]) ** 2 + (ax[None, :, None] - src[1]) ** 2 + (ax[None, None, :] - src[2]) ** 2
    far = r2.sqrt() > 10 * h
    out = {"grid": [n, n, n], "spacing_m": h, "source_m": src.tolist(),
           "nodes_excluded_within_m": 10 * h}
    v0, g = 1800.0, 2.7
    zc = (np.arange(n) + 0.5) * h
    cases = {"homogeneous_3000": (np.full((n, n, n), 3000.0),
                                  lambda: r2.sqrt() / 3000.0),
             "gradient_1800_plus_2.7z": (
                 np.ascontiguousarray(np.broadcast_to((v0 + g * zc)[:, None, None], (n, n, n))),
                 lambda: torch.acosh(1.0 + g * g * r2 / (2.0 * (v0 + g * src[0]) * (v0 + g * z))) / g)}
    for name, (v, exact) in cases.items():
        tt, _, _, sets = fa.Eikonal3D(v, (h, h, h)).solve_tensors(src, nsweep=None)
        t = exact()
        err = (tt[0] - t).abs()[far]
        rel = err / t[far]
        out[name] = {"sets": int(sets[0]), "max_abs_s": float(err.max()),
                     "max_rel": float(rel.max()), "mean_rel": float(rel.mean())}
        del tt
    log("analytic errors")
    return out


def fwd_adjoint(args, torch, fa, synthetic, dev):
    """Gradient workload: FP64 slowness 128 x 256 x 256, d = 10 m, true model smooth random
    (+-8% around 1/3000 s/m), gradient at its 1D average; 128 surface sources x 1024
    receivers; loss 0.5 sum (T - T_obs)^2.  One iteration = converged forward of all
    sources (with its sweep history), receiver interpolation, loss, exact adjoint, dL/ds."""
    from fteikpy_amd import _adjoint

    shape = tuple(int(s) for s in args.adj_shape.split(","))
    true, start, d, src, rec = synthetic.model_3d_smooth_slowness(shape=shape)
    st = torch.tensor(true, device=dev)
    s0 = torch.tensor(start, device=dev)
    tune = None
    if not args.no_tune:
        # the logged forward has the sweep structure of the plain one: tune on the plain
        tune = fa.autotune.tune3d(fa.Eikonal3D(1.0 / start, d), src)
    with torch.no_grad():
        tobs = fa.receiver_traveltimes(st, d, src, rec)
    del st

    def iteration(stats=None):
        s = s0.clone().requires_grad_(True)
        T = fa.receiver_traveltimes(s, d, src, rec, stats=stats)
        loss = 0.5 * ((T - tobs) ** 2).sum()
        loss.backward()
        return s.grad

    iteration()
    torch.cuda.synchronize()
    times, stats = [], {}
    for _ in range(args.adj_steps):
        t0 = time.perf_counter()
        g = iteration(stats)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    # the plain forward (no history) of the same sources, for the ratio
    rec_dev = torch.as_tensor(rec, device=dev)
    fwd = []
    for _ in range(2):
        t0 = time.perf_counter()
        tt, vz, _, _, _ = _adjoint._solve(s0, d, src, 200, False)
        _adjoint._interp(tt, d, np.zeros(3), src, vz, rec_dev)
        torch.cuda.synchronize()
        fwd.append(time.perf_counter() - t0)
        del tt
    log(f"fwd+adjoint: {args.adj_steps} iterations")
    s_iter = float(np.median(times))
    f = float(min(fwd))
    nodes = np.prod(np.array(shape) + 1) * len(src)
    return {"s_per_iter": s_iter, "iters": args.adj_steps, "forward_s": f,
            "ratio_to_forward": s_iter / f, "mode": "exact",
            "sets_per_source": [int(stats["sets"].min()), int(stats["sets"].max())],
            "history_entries_per_node": stats["events"] / nodes,
            "grad_finite": bool(torch.isfinite(g).all()),
            "autotune": tune,
            "config": {"workload": "adjoint-state gradient, smooth random slowness (+-8%), "
                                   "gradient at its 1D average",
                       "grid": list(shape), "spacing_m": d[0], "sources": len(src),
                       "receivers": len(rec), "loss": "0.5 sum (T - T_obs)^2"}}


if __name__ == "__main__":
    main()
