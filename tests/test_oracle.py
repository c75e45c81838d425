"""The oracle is pinned before anything is compared against it.

Checks, in order of strength:
- golden vectors written by the reference itself (tests/golden/gen_golden.py), on
  heterogeneous media, at nsweep = 1, 2, 3 and converged;
- every value the reference's own tests hold (tests/test_*.py of fteikpy 2.4.0);
- mathematics: closed forms in homogeneous media, order invariance of a directional
  pass, fixed-point idempotence, monotonicity across sets.
"""
import os

import numpy as np
import pytest

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _gold(dim):
    return np.load(os.path.join(GOLD, f"golden_{dim}d.npz"))


def _tags(g, s):
    k = int(g[f"s{s}_kconv"])
    return [(ns, f"s{s}_n{ns}" if ns != k else f"s{s}_conv") for ns in sorted({1, 2, 3, k})]


# ----------------------------------------------------------------------------------
# golden vectors from the reference (pure-Python mode)
# ----------------------------------------------------------------------------------
@pytest.mark.parametrize("dim", [2, 3])
def test_solve_matches_reference_golden(dim):
    g = _gold(dim)
    v, d, o = g["v"], g["gridsize"], g["origin"]
    solve = oracle.solve3d if dim == 3 else oracle.solve2d
    for s, src in enumerate(g["sources"]):
        for ns, tag in _tags(g, s):
            tt, gr, vz, _ = solve(1.0 / v, *d, src - o, nsweep=ns, grad=True)
            # traveltimes: bitwise on these fixtures (the reference's pow() and the
            # oracle's sqrt/x*x agree here); allow 1 ulp-level slack in the bound
            np.testing.assert_allclose(tt, g[tag + "_tt"], rtol=1e-15, atol=0)
            np.testing.assert_allclose(gr, g[tag + "_grad"], rtol=1e-13, atol=1e-15)
            assert vz == float(g[tag + "_vzero"])


@pytest.mark.parametrize("dim", [2, 3])
def test_convergence_count_matches_reference(dim):
    g = _gold(dim)
    v, d, o = g["v"], g["gridsize"], g["origin"]
    solve = oracle.solve3d if dim == 3 else oracle.solve2d
    for s, src in enumerate(g["sources"]):
        tt, _, _, k = solve(1.0 / v, *d, src - o, nsweep=0, max_sets=50)
        assert k == int(g[f"s{s}_kconv"])
        assert np.array_equal(tt, g[f"s{s}_conv_tt"])


@pytest.mark.parametrize("dim", [2, 3])
def test_batch_matches_reference_golden(dim):
    g = _gold(dim)
    v, d, o = g["v"], g["gridsize"], g["origin"]
    solve = oracle.solve3d if dim == 3 else oracle.solve2d
    tt, _, _, _ = solve(1.0 / v, *d, g["sources"][:2] - o, nsweep=2)
    np.testing.assert_allclose(tt, g["batch_tt"], rtol=1e-15, atol=0)


@pytest.mark.parametrize("dim", [2, 3])
def test_interp_and_rays_match_reference_golden(dim):
    g = _gold(dim)
    v, d, o = g["v"], g["gridsize"], g["origin"]
    axes_c = [o[a] + d[a] * np.arange(v.shape[a]) for a in range(dim)]
    interp = oracle.interp3d if dim == 3 else oracle.interp2d
    vinterp = oracle.vinterp3d if dim == 3 else oracle.vinterp2d
    ray = oracle.ray3d if dim == 3 else oracle.ray2d
    solve = oracle.solve3d if dim == 3 else oracle.solve2d
    np.testing.assert_array_equal(interp(*axes_c, v, g["vq"]), g["v_interp"])
    for s, src in enumerate(g["sources"]):
        tt, gr, vz, _ = solve(1.0 / v, *d, src - o, nsweep=int(g[f"s{s}_kconv"]), grad=True)
        axes = [o[a] + d[a] * np.arange(tt.shape[a]) for a in range(dim)]
        np.testing.assert_allclose(vinterp(*axes, tt, g[f"s{s}_q"], src, vz), g[f"s{s}_vinterp"],
                                   rtol=1e-14, atol=0)
        np.testing.assert_allclose(interp(*axes, np.ascontiguousarray(gr[..., 0]), g[f"s{s}_q"]),
                                   g[f"s{s}_gz_interp"], rtol=1e-12, atol=1e-14)
        stepsize = d.min()
        max_step = int(2.0 * sum((v.shape[a] * d[a]) ** 2 for a in range(dim)) ** 0.5 / stepsize)
        comps = [np.ascontiguousarray(gr[..., a]) for a in range(dim)]
        for hg in (0, 1):
            rays = ray(*axes, *comps, g[f"s{s}_rec"], src, stepsize, max_step, bool(hg))
            assert [len(r) for r in rays] == list(g[f"s{s}_rays_hg{hg}_len"])
            np.testing.assert_allclose(np.concatenate(rays), g[f"s{s}_rays_hg{hg}"], rtol=1e-12, atol=1e-9)


# ----------------------------------------------------------------------------------
# values from the reference's own tests (homogeneous, unit spacing)
# ----------------------------------------------------------------------------------
ONES2 = np.ones((16, 16))
ONES3 = np.ones((16, 16, 16))


@pytest.mark.parametrize("src, ref", [([0.0, 0.0], 3573.41870888), ([3.5, 3.5], 2462.96106736)])
def test_reference_pin_solve2d(src, ref):  # tests/test_eikonal2d.py:10-21
    tt = oracle.solve2d(ONES2, 1.0, 1.0, np.array(src), nsweep=2)[0]
    assert np.allclose(tt.sum(), ref, atol=1e-8)


@pytest.mark.parametrize("src, ref", [([0.0, 0.0, 0.0], 76631.78464590), ([3.5, 3.5, 3.5], 54091.60683054)])
def test_reference_pin_solve3d(src, ref):  # tests/test_eikonal3d.py:10-21
    tt = oracle.solve3d(ONES3, 1.0, 1.0, 1.0, np.array(src), nsweep=3)[0]
    assert np.allclose(tt.sum(), ref, atol=1e-8)


def test_reference_pin_meshio_sums():  # tests/test_meshio.py:9-24, :36-53
    tt, gr, _, _ = oracle.solve2d(np.ones((8, 10)), 1.0, 1.0, np.array([4.0, 5.0]), grad=True)
    assert np.allclose(tt.sum(), 378.23469225)
    assert tt[4, 5] == 0.0 and (gr[4, 5] == 0.0).all()
    assert np.allclose(gr.reshape(-1, 2).sum(0), 0.0)
    tt, gr, _, _ = oracle.solve3d(np.ones((8, 10, 12)), 1.0, 1.0, 1.0, np.array([4.0, 5.0, 6.0]), grad=True)
    assert np.allclose(tt.sum(), 6909.90160991)
    assert tt[4, 5, 6] == 0.0 and (gr[4, 5, 6] == 0.0).all()
    assert np.allclose(gr.reshape(-1, 3).sum(0), 0.0)


def _axes(shape):
    return [np.arange(n, dtype=np.float64) for n in shape]


@pytest.mark.parametrize("q, ref", [([0.0, 0.0], 0.0), ([0.5, 0.5], 0.70710678), ([1.5, 1.5], 2.12132034),
                                    ([15.0, 0.0], 15.0), ([15.0, 15.0], 21.21320343), ([0.0, 15.0], 15.0)])
def test_reference_pin_traveltime2d(q, ref):  # tests/test_traveltime2d.py:6-23
    tt, _, vz, _ = oracle.solve2d(ONES2, 1.0, 1.0, np.zeros(2), nsweep=2)
    assert np.allclose(oracle.vinterp2d(*_axes(tt.shape), tt, np.array(q), np.zeros(2), vz), ref, atol=1e-8)


@pytest.mark.parametrize("q, ref", [([0.0, 0.0, 0.0], 0.0), ([0.5, 0.5, 0.5], 0.86602540),
                                    ([1.5, 1.5, 1.5], 2.60631816), ([15.0, 15.0, 15.0], 25.98076211),
                                    ([15.0, 15.0, 0.0], 21.21320343)])
def test_reference_pin_traveltime3d(q, ref):  # tests/test_traveltime3d.py:6-32
    tt, _, vz, _ = oracle.solve3d(ONES3, 1.0, 1.0, 1.0, np.zeros(3), nsweep=3)
    assert np.allclose(oracle.vinterp3d(*_axes(tt.shape), tt, np.array(q), np.zeros(3), vz), ref, atol=1e-8)


def test_reference_pin_out_of_bound_is_nan():
    tt, _, vz, _ = oracle.solve3d(ONES3, 1.0, 1.0, 1.0, np.zeros(3), nsweep=3)
    assert np.isnan(oracle.vinterp3d(*_axes(tt.shape), tt, -np.ones(3), np.zeros(3), vz))


@pytest.mark.parametrize("p, ref2, ref3, hg", [(0.5, 1.0, 1.5, False), (0.5, 1.0, 1.5, True),
                                               (1.5, 4.75735931, 8.30384757, False), (1.5, 5.0, 7.5, True)])
def test_reference_pin_rays(p, ref2, ref3, hg):  # tests/test_traveltime{2,3}d.py:26-51
    tt, gr, _, _ = oracle.solve2d(ONES2, 1.0, 1.0, np.zeros(2), nsweep=2, grad=True)
    ray = oracle.ray2d(*_axes(tt.shape), gr[..., 0].copy(), gr[..., 1].copy(), np.full(2, p), np.zeros(2),
                       1.0, 46, hg)
    assert np.allclose(ray.sum(), ref2)
    tt, gr, _, _ = oracle.solve3d(ONES3, 1.0, 1.0, 1.0, np.zeros(3), nsweep=3, grad=True)
    ray = oracle.ray3d(*_axes(tt.shape), *[gr[..., a].copy() for a in range(3)], np.full(3, p), np.zeros(3),
                       1.0, 55, hg)
    assert np.allclose(ray.sum(), ref3)


def test_reference_pin_ray_lengths():  # tests/test_meshio.py:25-30, :55-59 (9- and 14-point rays)
    tt, gr, _, _ = oracle.solve2d(np.ones((8, 10)), 1.0, 1.0, np.array([4.0, 5.0]), grad=True)
    ray = oracle.ray2d(*_axes(tt.shape), gr[..., 0].copy(), gr[..., 1].copy(), np.zeros(2),
                       np.array([4.0, 5.0]), 1.0, 25, True)
    assert len(ray) == 9
    tt, gr, _, _ = oracle.solve3d(np.ones((8, 10, 12)), 1.0, 1.0, 1.0, np.array([4.0, 5.0, 6.0]), grad=True)
    ray = oracle.ray3d(*_axes(tt.shape), *[gr[..., a].copy() for a in range(3)], np.zeros(3),
                       np.array([4.0, 5.0, 6.0]), 1.0, 35, True)
    assert len(ray) == 14


def test_reference_pin_gradients():  # tests/test_gradient{2,3}d.py
    tt, gr, _, _ = oracle.solve3d(ONES3, 1.0, 1.0, 1.0, np.zeros(3), nsweep=3, grad=True)
    assert np.allclose(gr[15, 15, 15], 0.57735027) and np.allclose(gr[15, 0, 0], [1, 0, 0])
    assert np.allclose(gr[15, 0, 15], [0.70710678, 0, 0.70710678])
    tt, gr, _, _ = oracle.solve2d(ONES2, 1.0, 1.0, np.array([3.5, 3.5]), nsweep=2, grad=True)
    assert np.allclose(gr[0, 0], -0.70710678)


def test_errors():
    with pytest.raises(ValueError, match="source out of bound"):
        oracle.solve3d(ONES3, 1.0, 1.0, 1.0, np.array([-1.0, 0.0, 0.0]))
    with pytest.raises(ValueError, match="source out of bound"):
        oracle.solve2d(ONES2, 1.0, 1.0, np.array([0.0, 16.5]))
    tt, gr, _, _ = oracle.solve3d(ONES3, 1.0, 1.0, 1.0, np.zeros(3), nsweep=3, grad=True)
    comps = [gr[..., a].copy() for a in range(3)]
    with pytest.raises(ValueError, match="end point out of bound"):
        oracle.ray3d(*_axes(tt.shape), *comps, np.full(3, 20.0), np.zeros(3), 1.0, 50)
    with pytest.raises(RuntimeError, match="maximum number of steps"):
        oracle.ray3d(*_axes(tt.shape), *comps, np.full(3, 10.0), np.zeros(3), 1.0, 2)


# ----------------------------------------------------------------------------------
# mathematics
# ----------------------------------------------------------------------------------
def test_homogeneous_closed_forms():
    """On-node source, homogeneous medium: axis-aligned and grid-diagonal nodes are exact."""
    d = (2.0, 3.0, 4.0)
    tt = oracle.solve3d(np.full((9, 9, 9), 0.5), *d, np.array([8.0, 9.0, 12.0]), nsweep=3)[0]
    i0, j0, k0 = 4, 3, 3
    for n in range(5):  # +z axis
        assert abs(tt[i0 + n, j0, k0] - 0.5 * d[0] * n) < 1e-12
    for n in range(6):
        assert abs(tt[i0, j0 + n, k0] - 0.5 * d[1] * n) < 1e-12
    tt2 = oracle.solve2d(np.full((12, 12), 0.25), 1.0, 1.0, np.array([0.0, 0.0]), nsweep=2)[0]
    for n in range(13):
        assert abs(tt2[n, n] - 0.25 * n * 2 ** 0.5) < 1e-12
        assert abs(tt2[0, n] - 0.25 * n) < 1e-12


def test_pass_order_invariance_3d():
    """Within one directional pass every read is strictly upwind, so the i,j,k nesting
    gives the same state as the reference's k,j,i nesting bit for bit."""
    g = _gold(3)
    v, d, o = g["v"], g["gridsize"], g["origin"]
    for src in g["sources"]:
        for ns in (1, 3):
            a = oracle.solve3d(1.0 / v, *d, src - o, nsweep=ns, grad=True, order=0)
            b = oracle.solve3d(1.0 / v, *d, src - o, nsweep=ns, grad=True, order=1)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_fixed_point_idempotent_and_monotone():
    g = _gold(3)
    v, d, o = g["v"], g["gridsize"], g["origin"]
    src = g["sources"][0] - o
    prev = None
    for ns in range(1, 6):
        tt = oracle.solve3d(1.0 / v, *d, src, nsweep=ns)[0]
        if prev is not None:
            assert (tt <= prev).all()  # a set never increases a traveltime
        prev = tt
    conv, _, _, k = oracle.solve3d(1.0 / v, *d, src, nsweep=0)
    again = oracle.solve3d(1.0 / v, *d, src, nsweep=k + 3)[0]
    assert np.array_equal(conv, again)


def test_planted_mistake_is_caught():
    """A wrong spacing order (dz and dx swapped) must fail the golden comparison: the
    fixtures are sensitive to axis bookkeeping."""
    g = _gold(3)
    v, d, o = g["v"], g["gridsize"], g["origin"]
    src = g["sources"][0] - o
    tt = oracle.solve3d(1.0 / v, d[1], d[0], d[2], src[[1, 0, 2]] * 0 + src, nsweep=1)[0]
    assert not np.allclose(tt, g["s0_n1_tt"], rtol=1e-9)
