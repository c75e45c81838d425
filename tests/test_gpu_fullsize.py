"""GPU parity at the full sizes of the batched workloads.

- Many sources on one model (160 x 320 x 320, 12.5 m, 20 inclusions, 256 surface
  sources): all 256 solved at nsweep 1 and 2, four of them (corners and centre of the source
  grid) bit-identical to the oracle, with their 4096 receivers, and all 256 bit-identical
  on the persistent schedule; for every source of the
  converged batch, the grid one set before convergence already final (the last set
  changed nothing) and, for the four, two sets before not yet.
- Uneven batch of 48 models (16 2D up to 1024 per side, 32 3D up to 192 x 256 x 256, 765
  sources): the 16 2D models converged, bit-identical to the oracle with their set counts;
  the 3D models at nsweep 2 against the oracle on the first source of every model; for
  every 3D source of the converged batch, the grid one set before convergence already
  final.
- The gradient workload (128 x 256 x 256, 128 sources x 1024 receivers, starting model):
  Euler's identity for the degree-1 homogeneous solver, sum_c s_c dL/ds_c =
  sum_r (T_r - T_obs,r) T_r, on the benched gradient.

The oracle runs with its i,j,k loop order (order=1): the same node DAG as the reference's
k,j,i nesting, bit-identical results (tests/test_oracle.py pins the equivalence), 3-4x
faster on these grids.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from fteikpy_amd import synthetic

pytestmark = pytest.mark.gpu

PICK = [0, 15, 119, 255]  # source-grid corners and a centre source


@pytest.fixture(scope="module")
def fa(device):
    import fteikpy_amd

    return fteikpy_amd


def _bits(t):
    import torch

    return t.contiguous().view(torch.int64)


def _free():
    import torch

    torch.cuda.synchronize()
    torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def inclusions(fa):
    v, d, src, rec = synthetic.model_3d_inclusions()
    return v, d, src, rec, fa.Eikonal3D(v, d)


@pytest.mark.parametrize("nsweep", [1, 2])
def test_many_sources_first_sets_bitwise(fa, inclusions, nsweep):
    v, d, src, rec, eik = inclusions
    import torch

    print(f"160x320x320, 256 sources, nsweep {nsweep}: GPU", flush=True)
    tt, _, vz, sets = eik.solve_tensors(src, nsweep=nsweep)
    assert (sets == nsweep).all()
    got = tt[PICK].cpu().numpy()
    print("the persistent schedule, every source", flush=True)
    tp, _, vzp, setsp = eik.solve_tensors(src, nsweep=nsweep, schedule="persistent")
    assert torch.equal(_bits(tp), _bits(tt))
    assert np.array_equal(vzp, vz) and np.array_equal(setsp, sets)
    del tt, tp
    _free()
    print(f"oracle, sources {PICK}", flush=True)
    ref, _, rvz, _ = oracle.solve3d(1.0 / v, *d, src[PICK], nsweep=nsweep, order=1,
                                    nthreads=len(PICK))
    for k, s in enumerate(PICK):
        assert np.array_equal(got[k], ref[k]), s
        assert vz[s] == rvz[k]
    axes = [np.arange(n + 1) * di for n, di in zip(v.shape, d)]
    for k, s in enumerate(PICK):
        grid = fa.TraveltimeGrid3D(got[k], d, np.zeros(3), src[s], None, vz[s])
        np.testing.assert_array_equal(grid(rec),
                                      oracle.vinterp3d(*axes, ref[k], rec, src[s], rvz[k]))


def _final_set_changed_nothing(eik, src, conv, sets, label):
    """Sources whose converged solve took K sets: the same sources at nsweep = K - 1 give
    the converged grid bit for bit, i.e. the K-th set, which the solve ran, changed no
    node.  (A source retires after a set that changes nothing, so nsweep beyond K returns
    the converged grid by construction.)"""
    import torch

    for ks in np.unique(sets):
        idx = np.nonzero(sets == ks)[0]
        one, _, _, s1 = eik.solve_tensors(src[idx], nsweep=int(ks) - 1)
        assert (s1 == ks - 1).all(), (label, ks)
        for j, s in enumerate(idx):
            assert torch.equal(_bits(one[j]), _bits(conv[s])), (label, int(s), int(ks))
        del one


def test_many_sources_converged_fixed_point(fa, inclusions):
    v, d, src, _, eik = inclusions
    print("160x320x320, 256 sources, converged", flush=True)
    conv, _, _, sets = eik.solve_tensors(src, nsweep=None)
    print(f"sets per source {int(sets.min())}..{int(sets.max())}", flush=True)
    assert 2 < sets.min() and sets.max() < 100
    _final_set_changed_nothing(eik, src, conv, sets, "inclusions")
    for s in PICK:
        two, _, _, _ = eik.solve_tensors(src[s], nsweep=int(sets[s]) - 2)
        assert bool((two[0] >= conv[s]).all()) and bool((two[0] > conv[s]).any()), s
        del two
    del conv
    _free()


@pytest.fixture(scope="module")
def uneven(fa):
    models = synthetic.uneven_batch()
    probs = [((fa.Eikonal3D if v.ndim == 3 else fa.Eikonal2D)(v, d), src) for v, d, src in models]
    return models, probs


def test_uneven_batch_2d_models_bitwise(fa, uneven):
    models, probs = uneven
    print("uneven batch, converged", flush=True)
    res = fa.solve_batch_tensors(probs, nsweep=None)
    two_d = [(m, r) for m, r in zip(models, res) if m[0].ndim == 2]
    assert len(two_d) == 16
    got = [(tt.cpu().numpy(), vz, sets) for _, (tt, _, vz, sets) in two_d]
    del res
    _free()

    def ref(m):
        v, d, src = m
        return oracle.solve2d(1.0 / v, *d, src, nsweep=0)

    with ThreadPoolExecutor(8) as pool:
        refs = list(pool.map(ref, [m for m, _ in two_d]))
    for (m, _), (tt, vz, sets), (rtt, _, rvz, rk) in zip(two_d, got, refs):
        assert np.array_equal(tt, rtt), m[0].shape
        assert np.array_equal(vz, rvz) and np.array_equal(sets, rk), m[0].shape


def test_uneven_batch_3d_models(fa, uneven):
    models, probs = uneven
    three = [i for i, m in enumerate(models) if m[0].ndim == 3]
    assert len(three) == 32
    print("uneven batch, 3D models at nsweep 2", flush=True)
    res = fa.solve_batch_tensors([probs[i] for i in three], nsweep=2)
    got = [tt[0].cpu().numpy() for tt, _, _, _ in res]
    del res
    _free()

    def ref(i):
        v, d, src = models[i]
        return oracle.solve3d(1.0 / v, *d, src[0], nsweep=2, order=1)[0]

    with ThreadPoolExecutor(8) as pool:
        refs = list(pool.map(ref, three))
    for i, g, r in zip(three, got, refs):
        assert np.array_equal(g, r), models[i][0].shape

    print("uneven batch, 3D models converged; the last set of every source", flush=True)
    conv = fa.solve_batch_tensors([probs[i] for i in three], nsweep=None)
    for i, (tt, _, _, sets) in zip(three, conv):
        eik, src = probs[i]
        _final_set_changed_nothing(eik, np.asarray(src), tt, sets, models[i][0].shape)
    del conv
    _free()


def test_gradient_workload_euler_identity(fa):
    """T is homogeneous of degree 1 in the slowness (every candidate of the node update,
    and the seeded analytic times, scale with it), so its exact derivative satisfies
    sum_c s_c dT_r/ds_c = T_r; for L = 0.5 sum (T - T_obs)^2 this gives
    sum_c s_c dL/ds_c = sum_r (T_r - T_obs,r) T_r (a one-sided derivative at ties is the
    derivative of one branch, homogeneous as well)."""
    import torch

    true, start, d, src, rec = synthetic.model_3d_smooth_slowness()
    dev = "cuda"
    st = torch.tensor(true, device=dev)
    s0 = torch.tensor(start, device=dev, requires_grad=True)
    print("gradient workload: observed times", flush=True)
    with torch.no_grad():
        tobs = fa.receiver_traveltimes(st, d, src, rec)
    print("gradient workload: forward + adjoint at the starting model", flush=True)
    t = fa.receiver_traveltimes(s0, d, src, rec)
    loss = 0.5 * ((t - tobs) ** 2).sum()
    loss.backward()
    g = s0.grad
    assert torch.isfinite(g).all()
    lhs = float((s0.detach() * g).sum())
    rhs = float(((t.detach() - tobs) * t.detach()).sum())
    scale = float((s0.detach() * g).abs().sum())
    print(f"sum s dL/ds = {lhs:.12e}, sum (T - Tobs) T = {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) <= 1e-9 * scale
    del st, s0, tobs, t, g
    _free()
