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
    assert len(res["timings"]) >= 3
    assert autotune.lookup3d(shape, len(src)) == (res["tile_a"], res["group"])
    grids = eik.solve(src, nsweep=None)
    for s, g in zip(src, grids):
        ref, _, _, k = oracle.solve3d(1.0 / v, *d, s, nsweep=0)
        assert np.array_equal(g.grid, ref) and g.sets_done == k
    autotune.clear()
