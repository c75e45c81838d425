"""GPU: fteik3d_solve_noalloc, whose workspace is a file-scope device array, against
fteik3d_solve (workspace passed in) through the C ABI: outputs bit-identical on two calls
in a row, the second on a smaller problem (the array keeps the first call's contents),
both in converge mode and at a fixed nsweep, with and without gradient signs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rough(shape, seed):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(shape)
    for ax in range(3):
        p = (np.roll(p, 1, ax) + p + np.roll(p, -1, ax)) / 3.0
    return 1500.0 * np.exp(0.8 * p / np.abs(p).max())


def _run(noalloc, v, d, src, nsweep, grad):
    import torch

    from fteikpy_amd import _lib

    L = _lib.lib()
    nz, nx, ny = v.shape
    n = len(src)
    slow = torch.tensor(1.0 / v, device="cuda")
    tt = torch.full((n, nz + 1, nx + 1, ny + 1), -1.0, dtype=torch.float64, device="cuda")
    sgn = (torch.full(tt.shape, 0xAB, dtype=torch.uint8, device="cuda") if grad else None)
    vz = torch.zeros(n, dtype=torch.float64, device="cuda")
    sets = torch.zeros(n, dtype=torch.int32, device="cuda")
    s, sp = _lib.host_f8(np.asarray(src, dtype=np.float64))
    ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if noalloc:
        st = L.fteik3d_solve_noalloc(ptr(slow), nz, nx, ny, *d, sp, n, nsweep, 50, int(grad), 0,
                                     0, ptr(tt), ptr(sgn), ptr(vz), ptr(sets), stream)
    else:
        wsb = L.fteik3d_workspace_size(nz, nx, ny, n)
        ws = torch.empty(wsb, dtype=torch.uint8, device="cuda")
        st = L.fteik3d_solve(ptr(slow), nz, nx, ny, *d, sp, n, nsweep, 50, int(grad), 0, 0, 1,
                             ptr(tt), ptr(sgn), ptr(vz), ptr(sets), ptr(ws), wsb, 0, stream)
    assert st == 0, st
    torch.cuda.synchronize()
    out = [tt.cpu().numpy(), vz.cpu().numpy(), sets.cpu().numpy()]
    if grad:
        out.append(sgn.cpu().numpy())
    return out


CASES = [  # (shape, spacing, sources, nsweep, gradient signs)
    ((22, 19, 27), (7.0, 9.0, 8.0),
     [[3.3, 17.7, 40.1], [140.0, 60.0, 100.0], [147.0, 171.0, 216.0], [0.0, 0.0, 0.0]], 0, True),
    ((9, 12, 10), (10.0, 10.0, 10.0), [[45.0, 61.0, 33.3]], 3, False),
    # 40 sources: the default group runs its first two sets at 4 sources per workgroup
    ((14, 17, 12), (6.0, 5.0, 7.0),
     (np.random.default_rng(8).random((40, 3)) * np.array([84.0, 85.0, 84.0])).tolist(), 0,
     False),
]


@pytest.mark.parametrize("order", [0, 1])
def test_noalloc_matches_workspace_solve_twice(device, order):
    """Calls in a row, the second on the smaller problem (order 0) or on the larger one
    after the smaller (order 1), with a 40-source problem last or first."""
    from fteikpy_amd import _lib

    L = _lib.lib()
    cases = CASES if order == 0 else CASES[::-1]
    for k, (shape, d, src, nsweep, grad) in enumerate(cases):
        assert L.fteik3d_workspace_size(*shape, len(src)) <= L.fteik3d_scratch_capacity()
        v = _rough(shape, 31 + k)
        print(f"noalloc call {k}: shape {shape}, {len(src)} sources, nsweep {nsweep}", flush=True)
        ref = _run(False, v, d, src, nsweep, grad)
        got = _run(True, v, d, src, nsweep, grad)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)
        assert (ref[0] >= 0).all() and (ref[2] >= 1).all()
