"""GPU: the autotuner times real solves, records its choice, and the recorded choice is
used by later solves without changing a single bit of the result."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def test_tune3d_records_choice_and_result_is_unchanged(device):
    import fteikpy_amd
    from fteikpy_amd import autotune

    rng = np.random.default_rng(21)
    shape, d = (23, 30, 27), (10.0, 11.0, 9.0)
    v = 2000.0 + 2000.0 * rng.random(shape)
    src = rng.random((6, 3)) * np.array(shape) * d
    eik = fteikpy_amd.Eikonal3D(v, d)
    autotune.clear()
    res = autotune.tune3d(eik, src, tiles=(4, 8, 0), groups=(1, 2))
    print("tune3d", res, flush=True)
    assert res["tile_a"] in (0, 4, 8) and res["group"] in (1, 2)  # 0: the library default
    assert res["schedule"] in ("overlapped", "diagonal")
    assert len(res["timings"]) >= 4
    assert any(k.endswith(":diagonal") for k in res["timings"])
    assert autotune.lookup3d(shape, len(src)) == (res["tile_a"], res["group"])
    assert autotune.choice3d(shape, len(src)) == (res["tile_a"], res["group"], res["schedule"])
    grids = eik.solve(src, nsweep=None)
    for s, g in zip(src, grids):
        ref, _, _, k = oracle.solve3d(1.0 / v, *d, s, nsweep=0)
        assert np.array_equal(g.grid, ref) and g.sets_done == k
    autotune.clear()


def test_tune_batch_records_heights_and_result_is_unchanged(device):
    """tune_batch times whole-batch height rules and per-model heights, records one height
    per problem (0 for the 2D one), and solve_batch then uses them: every problem's times
    equal its own single solve bit for bit."""
    import fteikpy_amd
    from fteikpy_amd import autotune, synthetic

    models = synthetic.uneven_batch(seed=31, n2=1, n3=3, side2=(9, 30),
                                    side3=((8, 16), (8, 24)), nsrc=(1, 5))
    probs = [((fteikpy_amd.Eikonal3D if v.ndim == 3 else fteikpy_amd.Eikonal2D)(v, d), src)
             for v, d, src in models]
    autotune.clear()
    res = autotune.tune_batch(probs, heights=(0, 8, 12), refine=2)
    print("tune_batch", res, flush=True)
    assert len(res["tiles"]) == len(probs)
    assert all(t in (0, 8, 12) for t in res["tiles"])
    assert all(t == 0 for (eik, _), t in zip(probs, res["tiles"]) if len(eik.shape) == 2)
    assert len(res["timings"]) >= 3
    assert autotune.lookup_batch(probs) == res["tiles"]
    out = fteikpy_amd.solve_batch_tensors(probs, nsweep=None)
    for (eik, src), (tt, _, _, sets) in zip(probs, out):
        want = eik.solve(src, nsweep=None)
        want = want if isinstance(want, list) else [want]
        for i, w in enumerate(want):
            assert np.array_equal(tt[i].cpu().numpy(), w.grid)
            assert sets[i] == w.sets_done
    autotune.clear()
