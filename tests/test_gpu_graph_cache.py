"""GPU: sweep-set graphs kept across fteik3d_solve calls.  A later call with the same
buffers and problem replays the kept graph -- after a converging solve left it launching
fewer source groups -- and every call's traveltimes, signs and set counts equal those of
direct launches, at fixed nsweep and converged, on both schedules."""
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


class _Buffers:
    def __init__(self, torch, v, src, grad):
        nz, nx, ny = v.shape
        n = len(src)
        self.shape = v.shape
        self.n = n
        self.slow = torch.tensor(1.0 / v, device="cuda")
        self.tt = torch.empty((n, nz + 1, nx + 1, ny + 1), dtype=torch.float64, device="cuda")
        self.sgn = torch.empty(self.tt.shape, dtype=torch.uint8, device="cuda") if grad else None
        self.vz = torch.empty(n, dtype=torch.float64, device="cuda")
        self.sets = torch.empty(n, dtype=torch.int32, device="cuda")
        from fteikpy_amd import _lib

        self.wsb = _lib.lib().fteik3d_workspace_size(nz, nx, ny, n)
        self.ws = torch.empty(self.wsb, dtype=torch.uint8, device="cuda")
        self.src = np.asarray(src, dtype=np.float64)

    def solve(self, torch, stream, d, nsweep, graph, tile_a, group, schedule):
        from fteikpy_amd import _lib

        ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
        self.tt.fill_(-1.0)
        if self.sgn is not None:
            self.sgn.fill_(0xAB)
        self.sets.fill_(-1)
        torch.cuda.synchronize()
        _, sp = _lib.host_f8(self.src)
        nz, nx, ny = self.shape
        st = _lib.lib().fteik3d_solve(
            ptr(self.slow), nz, nx, ny, *d, sp, self.n, nsweep, 60,
            int(self.sgn is not None), tile_a, group, schedule, ptr(self.tt), ptr(self.sgn),
            ptr(self.vz), ptr(self.sets), ptr(self.ws), self.wsb, int(graph),
            ctypes.c_void_p(stream.cuda_stream))
        assert st == 0, st
        stream.synchronize()
        out = [self.tt.cpu().numpy(), self.sets.cpu().numpy()]
        if self.sgn is not None:
            out.append(self.sgn.cpu().numpy())
        return out


def _cache(clear=False):
    from fteikpy_amd import _lib

    h, m = ctypes.c_longlong(0), ctypes.c_longlong(0)
    held = _lib.lib().fteik3d_graph_cache(int(clear), ctypes.byref(h), ctypes.byref(m))
    return held, h.value, m.value


@pytest.mark.parametrize("tile_a,group,schedule,grad", [(0, 1, 2, False), (9, 1, 2, True),
                                                        (3, 2, 1, True)])
def test_kept_set_graph_replays_bitwise(device, tile_a, group, schedule, grad):
    import torch

    shape, d = (22, 19, 27), (7.0, 9.0, 8.0)
    v = _rough(shape, 41)
    # sources converging after different numbers of sets, so that converging solves launch
    # fewer source groups before they end
    src = [[3.3, 17.7, 40.1], [140.0, 60.0, 100.0], [147.0, 171.0, 216.0], [70.0, 85.5, 0.0],
           [0.0, 0.0, 0.0], [154.0, 171.0, 216.0]]
    buf = _Buffers(torch, v, src, grad)
    stream = torch.cuda.Stream()
    _cache(clear=True)
    ref = {ns: buf.solve(torch, stream, d, ns, False, tile_a, group, schedule) for ns in (0, 2)}
    assert len(set(ref[0][1].tolist())) > 1  # the sources converge after different sets
    held0, h0, m0 = _cache()
    assert held0 == 0
    order = (0, 0, 2, 0, 2)
    for k, ns in enumerate(order):
        print(f"graph call {k}: nsweep {ns}", flush=True)
        got = buf.solve(torch, stream, d, ns, True, tile_a, group, schedule)
        for a, b in zip(got, ref[ns]):
            assert np.array_equal(a, b), (k, ns)
    held, h, m = _cache()
    assert held == 1
    assert (h - h0, m - m0) == (len(order) - 1, 1)
    # another buffer set does not match the kept graph
    other = _Buffers(torch, v, src, grad)
    got = other.solve(torch, stream, d, 0, True, tile_a, group, schedule)
    for a, b in zip(got, ref[0]):
        assert np.array_equal(a, b)
    held, h2, m2 = _cache()
    assert held == 2 and (h2, m2) == (h, m + 1)
    assert _cache(clear=True)[0] == 0


def test_default_group_two_phases(device):
    """group 0 from 32 sources runs the first 2 sets at 4 sources per workgroup and the
    later ones at 2, each phase from its own kept graph: the same times, signs and set
    counts as groups 2 and 4 throughout, at nsweep 1-3 and converged, with and without
    graphs."""
    import torch

    shape, d = (22, 19, 27), (7.0, 9.0, 8.0)
    v = _rough(shape, 43)
    rng = np.random.default_rng(44)
    ext = np.array([shape[0] * d[0], shape[1] * d[1], shape[2] * d[2]])
    src = rng.random((40, 3)) * ext
    buf = _Buffers(torch, v, src, True)
    stream = torch.cuda.Stream()
    _cache(clear=True)
    for ns in (1, 2, 3, 0):
        print(f"two phases: nsweep {ns}", flush=True)
        ref = buf.solve(torch, stream, d, ns, False, 0, 2, 2)
        for group, graph in ((4, False), (0, False), (0, True), (0, True)):
            got = buf.solve(torch, stream, d, ns, graph, 0, group, 2)
            for a, b in zip(got, ref):
                assert np.array_equal(a, b), (ns, group, graph)
    held, h, m = _cache()
    # nsweep 1 captured the first phase only; nsweep 2 reused it and captured no other
    # (its second set is still in the first phase); nsweep 3 captured the second phase
    assert held == 2, held
    assert _cache(clear=True)[0] == 0
