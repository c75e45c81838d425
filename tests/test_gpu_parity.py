"""GPU parity: the HIP path against the oracle (itself pinned to the reference in
test_oracle.py) on the same inputs.

Bar: traveltimes are bit-identical to the oracle at every nsweep (each directional pass
is order-invariant and the library is built without FP contraction), and hence within
1e-15 of the reference's golden vectors; the number of sets to convergence matches;
gradients, interpolated values and ray points within 1e-12 relative (the required bar
for converged traveltimes is 1e-9, for interpolated values and ray points 1e-8).
"""
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def fa(device):
    import fteikpy_amd

    return fteikpy_amd


def _gold(dim):
    return np.load(os.path.join(GOLD, f"golden_{dim}d.npz"))


def _tags(g, s):
    k = int(g[f"s{s}_kconv"])
    return [(ns, f"s{s}_n{ns}" if ns != k else f"s{s}_conv") for ns in sorted({1, 2, 3, k})]


@pytest.mark.parametrize("dim", [3, 2])
def test_golden_solve(fa, dim):
    g = _gold(dim)
    v, d, o = g["v"], g["gridsize"], g["origin"]
    Eik = fa.Eikonal3D if dim == 3 else fa.Eikonal2D
    solve = oracle.solve3d if dim == 3 else oracle.solve2d
    eik = Eik(v, d, o)
    for s, src in enumerate(g["sources"]):
        print(f"golden {dim}D source {s}", flush=True)
        for ns, tag in _tags(g, s):
            tt = eik.solve(src, nsweep=ns, return_gradient=True)
            ref, rgrad, rvz, _ = solve(1.0 / v, *d, src - o, nsweep=ns, grad=True)
            assert np.array_equal(tt.grid, ref), (s, ns)
            np.testing.assert_allclose(tt.grid, g[tag + "_tt"], rtol=1e-15, atol=0)
            assert tt.vzero == rvz == float(g[tag + "_vzero"])
            grad = np.stack([c.grid for c in tt.gradient], -1)
            np.testing.assert_allclose(grad, rgrad, rtol=1e-13, atol=1e-15)
            np.testing.assert_allclose(grad, g[tag + "_grad"], rtol=1e-12, atol=1e-14)
        conv = eik.solve(src, nsweep=None)
        assert conv.sets_done == int(g[f"s{s}_kconv"])
        assert np.array_equal(conv.grid, g[f"s{s}_conv_tt"]) or np.allclose(
            conv.grid, g[f"s{s}_conv_tt"], rtol=1e-15, atol=0)


@pytest.mark.parametrize("dim", [3, 2])
def test_golden_interp_and_rays(fa, dim):
    g = _gold(dim)
    v, d, o = g["v"], g["gridsize"], g["origin"]
    Eik = fa.Eikonal3D if dim == 3 else fa.Eikonal2D
    eik = Eik(v, d, o)
    np.testing.assert_array_equal(eik(g["vq"]), g["v_interp"])
    for s, src in enumerate(g["sources"]):
        tt = eik.solve(src, nsweep=int(g[f"s{s}_kconv"]), return_gradient=True)
        np.testing.assert_allclose(tt(g[f"s{s}_q"]), g[f"s{s}_vinterp"], rtol=1e-13, atol=0)
        np.testing.assert_allclose(tt.gradient[0](g[f"s{s}_q"]), g[f"s{s}_gz_interp"], rtol=1e-12,
                                   atol=1e-14)
        for hg in (0, 1):
            rays = tt.raytrace(g[f"s{s}_rec"], honor_grid=bool(hg))
            assert [len(r) for r in rays] == list(g[f"s{s}_rays_hg{hg}_len"]), (s, hg)
            np.testing.assert_allclose(np.concatenate(rays), g[f"s{s}_rays_hg{hg}"], rtol=1e-12,
                                       atol=1e-9)


def test_batch_list_and_single_point_returns(fa):
    g = _gold(3)
    eik = fa.Eikonal3D(g["v"], g["gridsize"], g["origin"])
    tts = eik.solve(g["sources"][:2], nsweep=2)
    assert isinstance(tts, list) and len(tts) == 2
    np.testing.assert_allclose(np.stack([t.grid for t in tts]), g["batch_tt"], rtol=1e-15, atol=0)
    assert isinstance(tts[0](g["sources"][0]), float)
    assert np.isnan(tts[0](g["origin"] - 1.0))


def _rough_model(shape, seed):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(shape)
    for ax in range(len(shape)):
        p = (np.roll(p, 1, ax) + p + np.roll(p, -1, ax)) / 3.0
    return 1500.0 * np.exp(0.8 * p / np.abs(p).max())


@pytest.mark.parametrize("tile_a, group, schedule", [
    (1, 1, "diagonal"), (3, 2, "diagonal"), (8, 4, "diagonal"), (0, 0, "diagonal"),
    (64, 8, "diagonal"), (5, 4, "diagonal"),
    (8, 2, "overlapped"), (9, 1, "overlapped"), (0, 0, "overlapped"), (64, 8, "overlapped"),
    (10, 4, "overlapped"), (3, 2, "overlapped"),
    (3, 2, "persistent"), (0, 4, "persistent"), (1, 1, "persistent"), (9, 1, "persistent"),
    (64, 8, "persistent")])
def test_brick_shapes_bitwise(fa, tile_a, group, schedule):
    """Partial bricks in every axis, several z brick heights and source groups (groups
    partly empty), sources converging after different numbers of sets."""
    shape, d = (21, 19, 27), (7.0, 9.0, 8.0)
    v = _rough_model(shape, 11)
    srcs = np.array([[3.3, 17.7, 40.1], [140.0, 60.0, 100.0], [147.0, 171.0, 216.0],
                     [70.0, 85.5, 0.0], [0.0, 0.0, 0.0]])
    ref, _, _, kref = oracle.solve3d(1.0 / v, *d, srcs, nsweep=0)
    tt, _, _, sets = fa.Eikonal3D(v, d).solve_tensors(srcs, nsweep=None, tile_a=tile_a,
                                                      group=group, schedule=schedule)
    assert list(sets) == list(kref)
    assert np.array_equal(tt.cpu().numpy(), ref)


def test_graph_mode_matches_direct(fa):
    shape, d = (18, 22, 25), (10.0, 10.0, 10.0)
    v = _rough_model(shape, 3)
    srcs = np.array([[5.0, 5.0, 5.0], [170.0, 200.0, 240.0]])
    eik = fa.Eikonal3D(v, d)
    a = eik.solve_tensors(srcs, nsweep=None, use_graph=False)
    for _ in range(2):  # replays with eager work between them
        b = eik.solve_tensors(srcs, nsweep=None, use_graph=True)
        assert np.array_equal(a[0].cpu().numpy(), b[0].cpu().numpy())
        assert list(a[3]) == list(b[3])
        b2 = eik.solve_tensors(srcs, nsweep=3, use_graph=True)
        a2 = eik.solve_tensors(srcs, nsweep=3, use_graph=False)
        assert np.array_equal(a2[0].cpu().numpy(), b2[0].cpu().numpy())


def test_overlapped_equals_pass_by_pass(fa):
    """The overlapped schedule (bricks of several passes in one launch), one pass per
    launch and the persistent schedule (one cooperative launch per set) give identical
    traveltimes, upwind signs and set counts at every nsweep and converged, directly and
    from a captured graph, and equal the oracle."""
    shape, d = (40, 45, 38), (9.0, 10.0, 11.0)
    v = _rough_model(shape, 23)
    rng = np.random.default_rng(23)
    srcs = rng.uniform(0.0, 1.0, (6, 3)) * np.array(shape) * np.array(d)
    eik = fa.Eikonal3D(v, d)
    for ns in (1, 2, 3, None):
        for tile, group, graph in ((8, 2, False), (0, 0, True), (12, 1, True)):
            a = eik.solve_tensors(srcs, nsweep=ns, return_gradient=True, tile_a=tile,
                                  group=group, schedule="diagonal")
            b = eik.solve_tensors(srcs, nsweep=ns, return_gradient=True, tile_a=tile,
                                  group=group, schedule="overlapped", use_graph=graph)
            c = eik.solve_tensors(srcs, nsweep=ns, return_gradient=True, tile_a=tile,
                                  group=group, schedule="persistent")
            for x in (b, c):
                assert np.array_equal(a[0].cpu().numpy(), x[0].cpu().numpy()), (ns, tile)
                assert np.array_equal(a[1].cpu().numpy(), x[1].cpu().numpy()), (ns, tile)
                assert list(a[3]) == list(x[3]), (ns, tile)
    ref, _, _, kref = oracle.solve3d(1.0 / v, *d, srcs, nsweep=0)
    assert list(b[3]) == list(kref)
    assert np.array_equal(b[0].cpu().numpy(), ref)


def test_larger_model_bitwise(fa):
    shape, d = (48, 56, 64), (10.0, 10.0, 10.0)
    v = _rough_model(shape, 5)
    src = np.array([123.4, 301.7, 222.2])
    ref, rgrad, _, k = oracle.solve3d(1.0 / v, *d, src, nsweep=0, grad=True)
    tt = fa.Eikonal3D(v, d).solve(src, nsweep=None, return_gradient=True)
    assert tt.sets_done == k
    assert np.array_equal(tt.grid, ref)
    grad = np.stack([c.grid for c in tt.gradient], -1)
    np.testing.assert_allclose(grad, rgrad, rtol=1e-13, atol=1e-15)


def test_reference_pins(fa):
    """Values of the reference's own tests (tests/test_eikonal{2,3}d.py, test_traveltime*.py,
    test_gradient*.py)."""
    e3 = fa.Eikonal3D(np.ones((16, 16, 16)), (1.0, 1.0, 1.0))
    e2 = fa.Eikonal2D(np.ones((16, 16)), (1.0, 1.0))
    for src, ref in (([0.0, 0.0, 0.0], 76631.78464590), ([3.5, 3.5, 3.5], 54091.60683054)):
        assert np.allclose(e3.solve(src, nsweep=3).grid.sum(), ref)
    for src, ref in (([0.0, 0.0], 3573.41870888), ([3.5, 3.5], 2462.96106736)):
        assert np.allclose(e2.solve(src, nsweep=2).grid.sum(), ref)
    tt = e3.solve([0.0, 0.0, 0.0], nsweep=3, return_gradient=True)
    assert np.allclose(tt([1.5, 1.5, 1.5]), 2.60631816)
    assert np.allclose(tt([15.0, 15.0, 15.0]), 25.98076211)
    assert np.allclose(tt.gradient[0]([15.0, 15.0, 15.0]), 0.57735027)
    assert np.allclose(tt.raytrace([1.5, 1.5, 1.5], stepsize=1.0).sum(), 8.30384757)
    assert np.allclose(tt.raytrace([1.5, 1.5, 1.5], stepsize=1.0, honor_grid=True).sum(), 7.5)
    tt2 = e2.solve([0.0, 0.0], nsweep=2, return_gradient=True)
    assert np.allclose(tt2.raytrace([1.5, 1.5], stepsize=1.0).sum(), 4.75735931)
    assert np.allclose(tt2.raytrace([1.5, 1.5], stepsize=1.0, honor_grid=True).sum(), 5.0)
    assert np.allclose(tt2([0.5, 0.5]), 0.70710678)
    assert np.allclose(e2.solve([3.5, 3.5], nsweep=2, return_gradient=True).gradient[0]([0.0, 0.0]),
                       -0.70710678)
    # meshio-test sums and ray lengths (tests/test_meshio.py)
    t = fa.Eikonal3D(np.ones((8, 10, 12)), (1.0, 1.0, 1.0)).solve((4.0, 5.0, 6.0), return_gradient=True)
    assert np.allclose(t.grid.sum(), 6909.90160991)
    assert len(t.raytrace((0.0, 0.0, 0.0), honor_grid=True)) == 14
    t = fa.Eikonal2D(np.ones((8, 10)), (1.0, 1.0)).solve((4.0, 5.0), return_gradient=True)
    assert np.allclose(t.grid.sum(), 378.23469225)
    assert len(t.raytrace((0.0, 0.0), honor_grid=True)) == 9


def test_errors(fa):
    e3 = fa.Eikonal3D(np.ones((8, 8, 8)), (1.0, 1.0, 1.0))
    with pytest.raises(ValueError, match="source out of bound"):
        e3.solve([9.0, 0.0, 0.0])
    with pytest.raises(ValueError, match="source out of bound"):
        fa.Eikonal2D(np.ones((8, 8)), (1.0, 1.0)).solve([0.0, -0.1])
    tt = e3.solve([0.0, 0.0, 0.0], nsweep=2)
    with pytest.raises(ValueError, match="no gradient grid"):
        tt.gradient
    tt = e3.solve([0.0, 0.0, 0.0], nsweep=2, return_gradient=True)
    with pytest.raises(ValueError, match="end point out of bound"):
        tt.raytrace([9.5, 0.0, 0.0])
    with pytest.raises(RuntimeError, match="maximum number of steps"):
        tt.raytrace([8.0, 8.0, 8.0], stepsize=1.0, max_step=2)


def test_2d_many_sources(fa):
    g = _gold(2)
    v, d, o = g["v"], g["gridsize"], g["origin"]
    rng = np.random.default_rng(2)
    srcs = o + rng.random((37, 2)) * (np.array(v.shape) * d)
    ref = oracle.solve2d(1.0 / v, *d, srcs - o, nsweep=3)[0]
    tts = fa.Eikonal2D(v, d, o).solve(srcs, nsweep=3)
    assert np.array_equal(np.stack([t.grid for t in tts]), ref)


def test_2d_speculative_columns_bitwise(fa):
    """Tall rough 2D models: each z pass runs in blocks of 64 nodes evaluated speculatively
    and finished in order where a dependency chain is dense; 333 rows (blocks of 64 and a
    ragged 13), cell-scale contrast (branch switches, 3-point operators), sources in the
    interior, near a corner and on the last row; times, sets and gradients equal the
    oracle's at nsweep 1, 2, 3 and converged."""
    v = _rough_model((333, 150), 21)
    d = (7.0, 9.0)
    ext = np.array(v.shape) * d
    srcs = np.array([[0.37, 0.61], [0.02, 0.97], [1.0, 0.33], [0.5, 0.0]]) * ext
    eik = fa.Eikonal2D(v, d)
    for ns in (1, 2, 3, None):
        ref, rgrad, _, k = oracle.solve2d(1.0 / v, *d, srcs, nsweep=ns or 0, grad=True)
        tts = eik.solve(srcs, nsweep=ns, return_gradient=True)
        for s, t in enumerate(tts):
            assert np.array_equal(t.grid, ref[s]), (ns, s)
            g = np.stack([x.grid for x in t.gradient], -1)
            assert np.array_equal(g, rgrad[s]), (ns, s)
            if ns is None:
                assert t.sets_done == k[s], s


def test_2d_deeper_than_lds_bitwise(fa):
    """A model deeper than one workgroup's LDS holds (2300 cells: 2302 staged rows of 72 B
    exceed 160 KiB) stages its rows in global memory instead, with the same results; in a
    batch beside a model staged in LDS, too."""
    v = _rough_model((2300, 12), 22)
    d = (5.0, 6.0)
    srcs = np.array([[0.4, 0.5], [0.95, 0.1]]) * (np.array(v.shape) * d)
    for ns in (2, None):
        ref, _, _, k = oracle.solve2d(1.0 / v, *d, srcs, nsweep=ns or 0)
        tts = fa.Eikonal2D(v, d).solve(srcs, nsweep=ns)
        assert np.array_equal(np.stack([t.grid for t in tts]), ref), ns
    vs = _rough_model((50, 40), 23)
    ss = np.array([[0.3, 0.6]]) * (np.array(vs.shape) * d)
    out = fa.solve_batch_tensors([(fa.Eikonal2D(vs, d), ss), (fa.Eikonal2D(v, d), srcs)],
                                 nsweep=None)
    assert np.array_equal(out[0][0].cpu().numpy(), oracle.solve2d(1.0 / vs, *d, ss, nsweep=0)[0])
    assert np.array_equal(out[1][0].cpu().numpy(), oracle.solve2d(1.0 / v, *d, srcs, nsweep=0)[0])


@pytest.mark.parametrize("case", range(12))
def test_2d_random_models_bitwise(fa, case):
    """Speculative 2D passes on random models: row counts around the 64-node blocks (1, 2,
    63, 64, 65, 127, 128, 129 cells and others), smooth to cell-scale contrast, sources on
    edges and corners and in the interior; times, sets and (every other case) gradients
    equal the oracle's at nsweep 1, 2, 4 and converged."""
    rng = np.random.default_rng(500 + case)
    nz = [1, 2, 63, 64, 65, 127, 128, 129, 200, 31, 96, 257][case]
    nx = int(rng.integers(1, 90))
    base = _rough_model((nz, nx), 600 + case)
    contrast = rng.uniform(0.0, 1.0)
    v = 1500.0 + contrast * (base - 1500.0) + rng.uniform(0.0, 2000.0)
    d = tuple(rng.uniform(2.0, 20.0, 2))
    ext = np.array(v.shape) * d
    pos = rng.random((3, 2))
    pos[0] = rng.integers(0, 2, 2)  # a corner
    pos[1, rng.integers(0, 2)] = rng.integers(0, 2)  # an edge
    srcs = pos * ext
    grad = case % 2 == 0
    eik = fa.Eikonal2D(v, d)
    for ns in (1, 2, 4, None):
        ref, rgrad, _, k = oracle.solve2d(1.0 / v, *d, srcs, nsweep=ns or 0, grad=grad)
        tts = eik.solve(srcs, nsweep=ns, return_gradient=grad)
        for s, t in enumerate(tts):
            assert np.array_equal(t.grid, ref[s]), (case, ns, s)
            if grad:
                g = np.stack([x.grid for x in t.gradient], -1)
                assert np.array_equal(g, rgrad[s]), (case, ns, s)
            if ns is None:
                assert t.sets_done == k[s], (case, s)


def test_3d_random_models_bitwise(fa):
    """20 random 3D models (1 to 39 cells per side, smooth to cell-scale contrast, 1-5
    sources, some on a face) at random nsweep, brick height and sources per workgroup:
    times and set counts equal the oracle's; the converged ones again as one batch with
    the batch's default brick heights and group."""
    probs, refs = [], []
    for case in range(20):
        rng = np.random.default_rng(7000 + case)
        shape = tuple(int(x) for x in rng.integers(1, 40, 3))
        p = rng.standard_normal(shape)
        for _ in range(int(rng.integers(0, 3))):
            for ax in range(3):
                p = (np.roll(p, 1, ax) + p + np.roll(p, -1, ax)) / 3.0
        p = p / max(np.abs(p).max(), 1e-12)
        v = rng.uniform(1000, 4000) * np.exp(rng.uniform(0, 1.0) * p)
        d = tuple(rng.uniform(0.5, 30.0, 3))
        pos = rng.random((int(rng.integers(1, 6)), 3))
        if rng.random() < 0.5:
            pos[0, int(rng.integers(0, 3))] = float(rng.integers(0, 2))
        srcs = pos * (np.array(shape) * d)
        ns = [1, 2, 3, None][int(rng.integers(0, 4))]
        ref, _, _, k = oracle.solve3d(1.0 / v, *d, srcs, nsweep=ns or 0)
        eik = fa.Eikonal3D(v, d)
        tile, group = int(rng.choice([0, 1, 5, 8, 16, 27])), int(rng.choice([0, 1, 2, 4, 8]))
        tt, _, _, sets = eik.solve_tensors(srcs, nsweep=ns, tile_a=tile, group=group)
        assert np.array_equal(tt.cpu().numpy(), ref), (case, shape, ns, tile, group)
        if ns is None:
            assert np.array_equal(sets, k), case
            probs.append((eik, srcs))
            refs.append(ref)
    for (tt, _, _, _), ref in zip(fa.solve_batch_tensors(probs, nsweep=None), refs):
        assert np.array_equal(tt.cpu().numpy(), ref)


def test_3d_many_sources_default_group_bitwise(fa):
    """From 32 sources the default group runs the first two sweep sets at 4 sources per
    workgroup and the later ones at 2: 3 random models with 32-44 sources (some on faces
    and corners), at nsweep 1, 2, 3 and converged, on the overlapped and the pass-by-pass
    schedule: times and set counts equal the oracle's, gradient directions within the
    module's bar."""
    for case in range(3):
        rng = np.random.default_rng(7300 + case)
        shape = tuple(int(x) for x in rng.integers(12, 30, 3))
        p = rng.standard_normal(shape)
        for ax in range(3):
            p = (np.roll(p, 1, ax) + p + np.roll(p, -1, ax)) / 3.0
        v = 2500.0 * np.exp(0.6 * p / np.abs(p).max())
        d = tuple(rng.uniform(4.0, 12.0, 3))
        pos = rng.random((int(rng.integers(32, 45)), 3))
        pos[0] = 0.0
        pos[1, int(rng.integers(0, 3))] = 1.0
        srcs = pos * (np.array(shape) * d)
        eik = fa.Eikonal3D(v, d)
        for ns in (1, 2, 3, None):
            print(f"many sources: case {case} nsweep {ns}", flush=True)
            ref, rgrad, _, k = oracle.solve3d(1.0 / v, *d, srcs, nsweep=ns or 0, grad=True)
            for sched in ("overlapped", "diagonal"):
                tt, _, _, sets = eik.solve_tensors(srcs, nsweep=ns, group=0, tile_a=0,
                                                   schedule=sched)
                assert np.array_equal(tt.cpu().numpy(), ref), (case, ns, sched)
                if ns is None:
                    assert np.array_equal(sets, k), (case, sched)
            grids = eik.solve(srcs, nsweep=ns, return_gradient=True)
            for i, g in enumerate(grids):
                assert np.array_equal(g.grid, ref[i]), (case, ns, i)
                grad = np.stack([c.grid for c in g.gradient], -1)
                np.testing.assert_allclose(grad, rgrad[i], rtol=1e-13, atol=1e-15)


def test_3d_random_configurations_bitwise(fa):
    """24 random models of 20-55 cells per side with 1-64 sources, each at a random nsweep,
    brick height, group, schedule (overlapped, pass by pass, persistent), gradient flag and
    with or without graph replay: times and set counts equal the oracle's."""
    import torch

    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for case in range(24):
            rng = np.random.default_rng(91000 + case)
            shape = tuple(int(x) for x in rng.integers(20, 56, 3))
            p = rng.standard_normal(shape)
            for _ in range(int(rng.integers(0, 4))):
                for ax in range(3):
                    p = (np.roll(p, 1, ax) + p + np.roll(p, -1, ax)) / 3.0
            p = p / max(np.abs(p).max(), 1e-12)
            v = rng.uniform(1000, 5000) * np.exp(rng.uniform(0, 1.2) * p)
            d = tuple(rng.uniform(1.0, 20.0, 3))
            pos = rng.random((int(rng.choice([1, 3, 8, 17, 32, 40, 64])), 3))
            if rng.random() < 0.5:
                pos[0, int(rng.integers(0, 3))] = float(rng.integers(0, 2))
            srcs = pos * (np.array(shape) * d)
            ns = [1, 2, 3, None][int(rng.integers(0, 4))]
            tile = int(rng.choice([0, 0, 8, 12, 16, 27]))
            group = int(rng.choice([0, 0, 1, 2, 4]))
            sched = str(rng.choice(["overlapped", "overlapped", "diagonal", "persistent"]))
            grad = bool(rng.random() < 0.3)
            print(f"random configuration {case}: {shape} {len(srcs)} sources nsweep {ns} "
                  f"tile {tile} group {group} {sched} grad {grad}", flush=True)
            ref, _, _, k = oracle.solve3d(1.0 / v, *d, srcs, nsweep=ns or 0)
            eik = fa.Eikonal3D(v, d)
            tt, _, _, sets = eik.solve_tensors(srcs, nsweep=ns, tile_a=tile, group=group,
                                               schedule=sched, return_gradient=grad,
                                               use_graph=bool(case % 2))
            assert np.array_equal(tt.cpu().numpy(), ref), case
            if ns is None:
                assert np.array_equal(sets, k), case


def test_zero_velocity_block_bitwise(fa):
    """Zero-velocity cells (infinite slowness, no check in the reference either): candidates
    through them come out infinite, or NaN from the fixed-divisor division where the
    reference's division gives inf, and lose every min either way; times equal the
    oracle's bit for bit in 2D and 3D, at nsweep 2 and converged."""
    v2 = _rough_model((90, 70), 31)
    v2[30:45, 20:50] = 0.0
    d2 = (6.0, 8.0)
    s2 = np.array([[0.1, 0.1], [0.9, 0.8]]) * (np.array(v2.shape) * d2)
    v3 = _rough_model((20, 24, 22), 32)
    v3[8:12, 5:15, 6:16] = 0.0
    d3 = (9.0, 7.0, 8.0)
    s3 = np.array([[0.1, 0.1, 0.2], [0.9, 0.8, 0.7]]) * (np.array(v3.shape) * d3)
    with np.errstate(divide="ignore"):
        slow2, slow3 = 1.0 / v2, 1.0 / v3
    for ns in (2, None):
        ref = oracle.solve2d(slow2, *d2, s2, nsweep=ns or 0)[0]
        got = np.stack([t.grid for t in fa.Eikonal2D(v2, d2).solve(s2, nsweep=ns)])
        assert np.array_equal(got, ref, equal_nan=True), ("2D", ns)
        ref = oracle.solve3d(slow3, *d3, s3, nsweep=ns or 0)[0]
        got = np.stack([t.grid for t in fa.Eikonal3D(v3, d3).solve(s3, nsweep=ns)])
        assert np.array_equal(got, ref, equal_nan=True), ("3D", ns)


def test_many_sets_skip_bitwise(fa):
    """Inclusion model needing ~17 sets: late sets change <1% of the nodes, so most node
    updates are skipped by the change tracking; results stay bit-identical at every
    intermediate nsweep and at convergence, for several brick/group shapes."""
    from fteikpy_amd import synthetic

    v, d, src, _ = synthetic.model_3d_inclusions(shape=(40, 80, 80), nsrc_side=4)
    srcs = src[[0, 5, 10, 15]]
    eik = fa.Eikonal3D(v, d)
    ref, _, _, kref = oracle.solve3d(1.0 / v, *d, srcs, nsweep=0)
    assert kref.max() >= 10
    for tile, group, sched in ((0, 4, "diagonal"), (16, 1, "diagonal"), (7, 2, "diagonal"),
                               (5, 8, "diagonal"), (1, 1, "diagonal"), (0, 4, "overlapped"),
                               (8, 1, "overlapped"), (13, 2, "overlapped"),
                               (16, 8, "overlapped")):
        tt, _, _, sets = eik.solve_tensors(srcs, nsweep=None, tile_a=tile, group=group,
                                           schedule=sched, use_graph=True)
        assert list(sets) == list(kref), (tile, group)
        assert np.array_equal(tt.cpu().numpy(), ref), (tile, group)
    for ns in (5, 9):
        r = oracle.solve3d(1.0 / v, *d, srcs, nsweep=ns, grad=True)
        tt, sg, vz, _ = eik.solve_tensors(srcs, nsweep=ns, return_gradient=True)
        assert np.array_equal(tt.cpu().numpy(), r[0]), ns
        g = fa.TraveltimeGrid3D(tt[1], d, np.zeros(3), srcs[1], None, vz[1], sign=sg[1])
        np.testing.assert_allclose(np.stack([c.grid for c in g.gradient], -1), r[1][1],
                                   rtol=1e-13, atol=1e-15)


def test_default_brick_heights_bitwise(fa):
    """The library's default brick heights (27 planes for these 53 node planes from 32
    sources, leaving a partial last brick; 16 from 16 sources): every intermediate nsweep
    and the converged grids stay bit-identical to the oracle."""
    shape, d = (52, 27, 30), (9.0, 8.0, 10.0)
    v = _rough_model(shape, 17)
    rng = np.random.default_rng(17)
    eik = fa.Eikonal3D(v, d)
    for n in (40, 16):
        srcs = rng.random((n, 3)) * np.array(shape) * d
        ref, _, _, kref = oracle.solve3d(1.0 / v, *d, srcs, nsweep=0)
        assert kref.max() > kref.min()  # sources retire at different sets
        for use_graph in (True, False):
            tt, _, _, sets = eik.solve_tensors(srcs, nsweep=None, use_graph=use_graph)
            assert list(sets) == list(kref)
            assert np.array_equal(tt.cpu().numpy(), ref)
        for ns in (1, 2, 3):
            r = oracle.solve3d(1.0 / v, *d, srcs, nsweep=ns)[0]
            tt = eik.solve_tensors(srcs, nsweep=ns, use_graph=True)[0]
            assert np.array_equal(tt.cpu().numpy(), r)


@pytest.mark.parametrize("tile_a, group", [(8, 2), (16, 1), (0, 0)])
def test_first_set_skip_boundary_sources_bitwise(fa, tile_a, group):
    """The first set already skips nodes whose 7 upwind neighbours still hold 1e5, and tile
    faces carry the bytes of nodes a pass does not update (the grid's first plane in that
    direction): sources on the surface, on edges and in corners, where those planes are
    reached first, stay bit-identical after 1 and 2 sets and at convergence."""
    shape, d = (26, 21, 19), (8.0, 10.0, 9.0)
    v = _rough_model(shape, 23)
    v[:2] *= 0.4  # slow surface layer: head waves run along the first node planes
    ext = np.array(shape) * d
    srcs = np.array([[0.0, 101.3, 77.7], [0.0, 0.0, 0.0], [ext[0], ext[1], 0.0],
                     [113.0, 0.0, ext[2]], [0.0, ext[1], 45.5], [96.1, 117.0, 83.2]])
    eik = fa.Eikonal3D(v, d)
    for ns in (1, 2):
        r = oracle.solve3d(1.0 / v, *d, srcs, nsweep=ns)[0]
        tt = eik.solve_tensors(srcs, nsweep=ns, tile_a=tile_a, group=group)[0]
        assert np.array_equal(tt.cpu().numpy(), r), ns
    ref, _, _, kref = oracle.solve3d(1.0 / v, *d, srcs, nsweep=0)
    tt, _, _, sets = eik.solve_tensors(srcs, nsweep=None, tile_a=tile_a, group=group)
    assert list(sets) == list(kref)
    assert np.array_equal(tt.cpu().numpy(), ref)


def test_first_set_guard_tiny_slowness_bitwise(fa):
    """Slowness x spacing below 1e-9 s turns the first-set skipping off (candidates built
    from untouched neighbours could round below 1e5 there): results stay bit-identical."""
    shape, d = (17, 15, 20), (1.0, 1.0, 1.0)
    v = 1.0e12 * _rough_model(shape, 29) / 1500.0
    srcs = np.array([[3.2, 7.7, 11.1], [0.0, 15.0, 20.0]])
    eik = fa.Eikonal3D(v, d)
    for ns in (1, 0):
        ref, _, _, kref = oracle.solve3d(1.0 / v, *d, srcs, nsweep=ns)
        tt, _, _, sets = eik.solve_tensors(srcs, nsweep=ns or None, tile_a=8, group=2)
        assert np.array_equal(tt.cpu().numpy(), ref), ns
