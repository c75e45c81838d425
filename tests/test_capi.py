"""CPU checks of the C ABI: the library loads and exports every declared entry point, and
the host-side argument checks reject bad input without touching a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "fteik_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fteik\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from fteikpy_amd import _lib

    L = _lib.lib()
    names = _declared()
    assert names, "no declarations found"
    for name in names:
        assert hasattr(L, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert set(_lib.SIGNATURES) == set(names)
    assert b"gfx950" in L.fteik_version()


def test_workspace_sizes():
    from fteikpy_amd import _lib

    L = _lib.lib()
    assert L.fteik3d_workspace_size(4, 4, 4, 1) >= 125
    assert L.fteik3d_workspace_size(4, 4, 4, 256) > L.fteik3d_workspace_size(4, 4, 4, 1)
    assert L.fteik2d_workspace_size(4) > 0


def test_host_argument_checks():
    """Calls that fail validation return before any HIP call (no device needed)."""
    from fteikpy_amd import _lib

    L = _lib.lib()
    src = np.zeros(3)
    sp = src.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    # null slowness pointer
    st = L.fteik3d_solve(None, 4, 4, 4, 1.0, 1.0, 1.0, sp, 1, 2, 10, 0, 0, 0, 0, None, None, None,
                         None, None, 0, 0, None)
    assert st == _lib.ERR_ARG
    with pytest.raises(ValueError, match="invalid argument"):
        _lib.check(st)
    with pytest.raises(ValueError, match="source out of bound"):
        _lib.check(_lib.ERR_SOURCE_OOB)
    with pytest.raises(RuntimeError, match="maximum number of steps"):
        _lib.check(_lib.ERR_MAX_STEP)


def test_noalloc_capacity_checks_before_any_device_work():
    """fteik3d_solve_noalloc refuses problems beyond its fixed scratch array, and bad
    pointers, before any HIP call."""
    from fteikpy_amd import _lib

    L = _lib.lib()
    cap = L.fteik3d_scratch_capacity()
    assert cap == 64 << 20
    src = np.zeros(3)
    sp = src.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: both calls return before any use
    n = 1
    while L.fteik3d_workspace_size(99, 99, 99, n) <= cap:
        n += 1
    assert L.fteik3d_workspace_size(99, 99, 99, n - 1) <= cap
    st = L.fteik3d_solve_noalloc(fake, 99, 99, 99, 1.0, 1.0, 1.0, sp, n, 2, 10, 0, 0, 0, fake,
                                 None, fake, fake, None)
    assert st == _lib.ERR_WORKSPACE
    st = L.fteik3d_solve_noalloc(None, 9, 9, 9, 1.0, 1.0, 1.0, sp, 1, 2, 10, 0, 0, 0, fake,
                                 None, fake, fake, None)
    assert st == _lib.ERR_ARG
    st = L.fteik3d_solve_noalloc(fake, 9, 9, 9, 1.0, 1.0, 1.0, sp, 1, 2, 10, 1, 0, 0, fake,
                                 None, fake, fake, None)
    assert st == _lib.ERR_ARG  # signs requested without a sign buffer


def test_source_check_before_any_device_work():
    """An out-of-bound source is rejected from the host (the reference raises before
    allocating, fteikpy/_fteik/_fteik3d.py:402-407): fake non-null device pointers are
    never dereferenced."""
    from fteikpy_amd import _lib

    L = _lib.lib()
    fake = ctypes.c_void_p(0x1000)
    src = np.array([-1.0, 0.0, 0.0])
    sp = src.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ws = L.fteik3d_workspace_size(4, 4, 4, 1)
    st = L.fteik3d_solve(fake, 4, 4, 4, 1.0, 1.0, 1.0, sp, 1, 2, 10, 0, 0, 0, 1, fake, None, fake, fake,
                         fake, ws, 0, None)
    assert st == _lib.ERR_SOURCE_OOB
    src2 = np.array([0.0, 4.5])
    sp2 = src2.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    st = L.fteik2d_solve(fake, 4, 4, 1.0, 1.0, sp2, 1, 2, 10, 0, fake, None, fake, fake, fake,
                         L.fteik2d_workspace_size(1), None)
    assert st == _lib.ERR_SOURCE_OOB


def test_product_does_not_import_oracle():
    """The product package must not reach the checker."""
    pkg = os.path.join(ROOT, "fteikpy_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "from oracle" not in text, f
                assert "libfteik_oracle" not in text, f


def test_schedule_argument_checks():
    """Unknown schedules are refused, and the persistent schedule without a device able to
    run cooperative launches (none here), before any device work."""
    from fteikpy_amd import _lib

    L = _lib.lib()
    fake = ctypes.c_void_p(0x1000)
    src = np.array([1.0, 1.0, 1.0])
    sp = src.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ws = L.fteik3d_workspace_size(4, 4, 4, 1)
    st = L.fteik3d_solve(fake, 4, 4, 4, 1.0, 1.0, 1.0, sp, 1, 2, 10, 0, 0, 0, 0, fake, None, fake,
                         fake, fake, ws, 0, None)
    assert st == _lib.ERR_UNSUPPORTED
    for bad in (-1, 3):  # no such schedule
        st = L.fteik3d_solve(fake, 4, 4, 4, 1.0, 1.0, 1.0, sp, 1, 2, 10, 0, 0, 0, bad, fake, None,
                             fake, fake, fake, ws, 0, None)
        assert st == _lib.ERR_ARG


def test_adjoint_host_argument_checks():
    """Adjoint and logged-solve entry points validate before touching a device."""
    from fteikpy_amd import _lib

    L = _lib.lib()
    o = np.zeros(3)
    op = o.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert L.fteik3d_history_size(-1, 8) == 0
    assert L.fteik3d_history_size(100, 8) >= 100 * 12
    assert L.fteik3d_adjoint_history_workspace_size(4, 4, 4, 2, 10) > 2 * 125 * 8
    assert L.fteik3d_adjoint_workspace_size(4, 4, 4, 2) >= 2 * 125 * 21
    ev = ctypes.c_longlong(0)
    # no history buffer
    st = L.fteik3d_solve_logged(None, 4, 4, 4, 1.0, 1.0, 1.0, op, 1, 0, 10, 0, 0, None, None,
                                None, None, 0, None, 0, ctypes.byref(ev), 0, None)
    assert st == _lib.ERR_ARG
    # (source, node) indices beyond 32 bits
    st = L.fteik3d_solve_logged(None, 1000, 1000, 1000, 1.0, 1.0, 1.0, op, 8, 0, 10, 0, 0, None,
                                None, None, None, 0, ctypes.c_void_p(1), 1 << 20,
                                ctypes.byref(ev), 0, None)
    assert st == _lib.ERR_ARG
    st = L.fteik3d_adjoint_history(None, 4, 4, 4, 1.0, 1.0, 1.0, op, op, 1, None, None, None, 0,
                                   None, None, None, 0, None)
    assert st == _lib.ERR_ARG
    st = L.fteik3d_adjoint(None, 4, 4, 4, 1.0, 1.0, 1.0, op, op, 1, None, None, 0, None, None,
                           None, None, None, 0, None)
    assert st == _lib.ERR_ARG
    with pytest.raises(RuntimeError, match="do not reproduce"):
        _lib.check(_lib.ERR_UNSETTLED)


def test_receiver_traveltimes_refuses_host_tensors():
    import torch

    import fteikpy_amd

    s = torch.full((4, 4, 4), 1e-3, dtype=torch.float64)
    with pytest.raises(ValueError, match="HIP device"):
        fteikpy_amd.receiver_traveltimes(s, (1.0, 1.0, 1.0), [[1.0, 1.0, 1.0]], [[2.0, 2.0, 2.0]])
    with pytest.raises(TypeError):
        fteikpy_amd.receiver_traveltimes(s.float(), (1.0, 1.0, 1.0), [[1.0, 1.0, 1.0]],
                                         [[2.0, 2.0, 2.0]])


def test_history_size_round_trip():
    """A buffer of fteik3d_history_size(n, p) bytes holds exactly n changes (the library
    derives the capacity from the byte count)."""
    from fteikpy_amd import _lib

    L = _lib.lib()
    for n in (1, 16, 17, 1000, 123457):
        for p in (8, 400):
            assert L.fteik3d_history_size(n, p) < L.fteik3d_history_size(n + 64, p)


def _problem(ndim, shape, src, fake):
    from fteikpy_amd import _lib

    p = _lib.Problem()
    p.ndim = ndim
    p.nz, p.nx = shape[0], shape[1]
    p.ny = shape[2] if ndim == 3 else 0
    p.dz = p.dx = p.dy = 1.0
    p.slow = p.tt = p.vzero = p.sets_done = fake
    p.sgn = None
    p.src = src.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    p.nsrc = len(src)
    p.tile_a = 0
    return p


def test_batch_checks_every_source_before_device_work():
    """fteik_solve_batch checks all problems on the host first: an out-of-bound source in
    any model returns code 1 with that model's index, and nothing touches the fake device
    pointers."""
    from fteikpy_amd import _lib

    L = _lib.lib()
    fake = 0x1000
    good3 = np.array([[1.0, 2.0, 3.0], [4.0, 0.5, 0.5]])
    good2 = np.array([[1.0, 2.0]])
    bad2 = np.array([[1.0, 2.0], [0.0, 9.5]])
    arr = (_lib.Problem * 3)(_problem(3, (5, 6, 7), good3, fake),
                             _problem(2, (8, 9), good2, fake),
                             _problem(2, (8, 9), bad2, fake))
    wsb = L.fteik_batch_workspace_size(arr, 3)
    assert wsb > 0
    failed = ctypes.c_int(-7)
    st = L.fteik_solve_batch(arr, 3, 0, 10, 0, 0, ctypes.c_void_p(fake), wsb, 0,
                             ctypes.byref(failed), None)
    assert st == _lib.ERR_SOURCE_OOB and failed.value == 2
    # a 3D source out of bound is found as well
    bad3 = np.array([[1.0, 2.0, 7.5]])
    arr[0] = _problem(3, (5, 6, 7), bad3, fake)
    arr[2] = _problem(2, (8, 9), good2, fake)
    st = L.fteik_solve_batch(arr, 3, 0, 10, 0, 0, ctypes.c_void_p(fake), wsb, 0,
                             ctypes.byref(failed), None)
    assert st == _lib.ERR_SOURCE_OOB and failed.value == 0
    # bad ndim, too small a workspace, gradients without sign buffers
    arr[0] = _problem(4, (5, 6, 7), good3, fake)
    assert L.fteik_solve_batch(arr, 3, 0, 10, 0, 0, ctypes.c_void_p(fake), wsb, 0, None,
                               None) == _lib.ERR_ARG
    arr[0] = _problem(3, (5, 6, 7), good3, fake)
    assert L.fteik_solve_batch(arr, 3, 0, 10, 0, 0, ctypes.c_void_p(fake), wsb - 1, 0, None,
                               None) == _lib.ERR_WORKSPACE
    assert L.fteik_solve_batch(arr, 3, 0, 10, 1, 0, ctypes.c_void_p(fake), wsb, 0, None,
                               None) == _lib.ERR_ARG
    assert L.fteik_solve_batch(arr, 3, 0, 10, 0, 3, ctypes.c_void_p(fake), wsb, 0, None,
                               None) == _lib.ERR_ARG


def test_autotune_records_and_persists(tmp_path):
    from fteikpy_amd import autotune

    autotune.clear()
    assert autotune.lookup3d((10, 20, 30), 4) == (0, 0)
    autotune.record3d((10, 20, 30), 4, 16, 2)
    assert autotune.lookup3d((10, 20, 30), 4) == (16, 2)
    assert autotune.lookup3d((10, 20, 30), 5) == (0, 0)
    assert autotune.choice3d((10, 20, 30), 4) == (16, 2, None)
    autotune.record3d((10, 20, 30), 8, 24, 1, "diagonal")
    assert autotune.choice3d((10, 20, 30), 8) == (24, 1, "diagonal")
    autotune.record3d((10, 20, 30), 9, 24, 1, "persistent")
    assert autotune.choice3d((10, 20, 30), 9) == (24, 1, "persistent")
    with pytest.raises(ValueError):
        autotune.record3d((10, 20, 30), 8, 24, 1, "bogus")

    class _Eik:  # stand-ins: the batch key needs each problem's shape and source count
        def __init__(self, shape):
            self.shape = shape

    probs = [(_Eik((10, 20, 30)), np.zeros((3, 3))), (_Eik((40, 50)), np.zeros(2)),
             (_Eik((6, 7, 8)), np.zeros(3))]
    assert autotune.batch_signature(probs) == "3d10x20x30:3;2d40x50:1;3d6x7x8:1"
    assert autotune.lookup_batch(probs) is None
    autotune.record_batch(probs, [16, 0, 8])
    assert autotune.lookup_batch(probs) == [16, 0, 8]
    assert autotune.lookup_batch(probs[:2]) is None
    with pytest.raises(ValueError):
        autotune.record_batch(probs, [16, 0])
    path = str(tmp_path / "tune.json")
    autotune.save(path)
    autotune.clear()
    autotune.load(path)
    assert autotune.lookup3d((10, 20, 30), 4) == (16, 2)
    assert autotune.choice3d((10, 20, 30), 8) == (24, 1, "diagonal")
    assert autotune.lookup_batch(probs) == [16, 0, 8]
    autotune.clear()
    # a file of (tile_a, group) pairs from before schedules were tuned still loads
    old = tmp_path / "old.json"
    old.write_text('{"10x20x30:4": [16, 2]}')
    autotune.load(str(old))
    assert autotune.choice3d((10, 20, 30), 4) == (16, 2, None)
    autotune.clear()


def test_autotune_library_defaults():
    """default3d states the library's defaults: group 0 (two phases: 4 sources per
    workgroup in the first two sets, then 2) from 32 sources, 2 from 8, 1 below; the
    brick height of default_tile."""
    from fteikpy_amd import autotune

    assert autotune.default3d(256) == (0, 0)
    assert autotune.default3d(32) == (0, 0)
    assert autotune.default3d(31) == (0, 2)
    assert autotune.default3d(8) == (0, 2)
    assert autotune.default3d(7) == (0, 1)
    assert autotune.default3d(1) == (0, 1)
    assert autotune.default_tile(256, 161) == 27 and autotune.default_tile(128, 129) == 26
    assert autotune.default_tile(16, 161) == 16 and autotune.default_tile(1, 201) == 8


def test_cell_plane_limit():
    """In-plane cell offsets of the 3D sweep are 32-bit: a cell plane of more than 2^31 - 1
    cells is refused from the host, before any device work (fake pointers)."""
    from fteikpy_amd import _lib

    L = _lib.lib()
    fake = ctypes.c_void_p(0x1000)
    src = np.array([0.5, 0.5, 0.5])
    sp = src.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    st = L.fteik3d_solve(fake, 1, 50000, 50000, 1.0, 1.0, 1.0, sp, 1, 2, 10, 0, 0, 0, 1, fake,
                         None, fake, fake, fake, 1 << 62, 0, None)
    assert st == _lib.ERR_ARG


def test_spacing_range():
    """The 2D and 3D operators' fixed-denominator division needs normal inverse squares:
    spacings outside [1e-100, 1e100] are refused from the host (fake pointers, no device
    work)."""
    from fteikpy_amd import _lib

    L = _lib.lib()
    fake = ctypes.c_void_p(0x1000)
    src = np.array([0.5e-120, 0.5, 0.5])
    sp = src.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for d in ((1e-120, 1.0, 1.0), (1.0, 1e120, 1.0), (1.0, 1.0, 0.0)):
        st = L.fteik3d_solve(fake, 4, 4, 4, *d, sp, 1, 2, 10, 0, 0, 0, 1, fake, None, fake, fake,
                             fake, 1 << 40, 0, None)
        assert st == _lib.ERR_ARG, d
    src2 = np.array([0.5, 0.5])
    sp2 = src2.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for d in ((1e-120, 1.0), (1.0, 1e120), (float("nan"), 1.0)):
        st = L.fteik2d_solve(fake, 4, 4, *d, sp2, 1, 2, 10, 0, fake, None, fake, fake, fake,
                             1 << 40, None)
        assert st == _lib.ERR_ARG, d


def test_sweep_launches_per_set():
    """Sweep-kernel launches per sweep set (host only): with the overlapped schedule 508
    against 688 one pass at a time for 160 x 320 x 320 cells in the default 27-plane bricks
    of 256 sources, 368 against 608 for 200^3 cells in one source's 8-plane bricks; bricks
    under 8 planes run pass by pass; invalid arguments give 0."""
    from fteikpy_amd import _lib

    L = _lib.lib()
    assert L.fteik3d_sweep_launches(160, 320, 320, 256, 0, 2) == 508
    assert L.fteik3d_sweep_launches(160, 320, 320, 256, 0, 1) == 688
    assert L.fteik3d_sweep_launches(200, 200, 200, 1, 0, 2) == 368
    assert L.fteik3d_sweep_launches(200, 200, 200, 1, 0, 1) == 8 * (26 + 26 + 26 - 2)
    assert L.fteik3d_sweep_launches(200, 200, 200, 1, 4, 2) == 8 * (51 + 26 + 26 - 2)
    assert L.fteik3d_sweep_launches(160, 320, 320, 256, 0, 0) == 1
    for bad in ((0, 4, 4, 1, 0, 2), (4, 4, 4, 0, 0, 2), (4, 4, 4, 1, -1, 2), (4, 4, 4, 1, 0, 3)):
        assert L.fteik3d_sweep_launches(*bad) == 0


def test_raytrace_batch_argument_checks():
    """The batched ray entry point validates its arguments before touching a device."""
    from fteikpy_amd import _lib

    L = _lib.lib()
    fake = ctypes.c_void_p(0x1000)
    o, op = _lib.host_f8(np.zeros(3))
    d, dp = _lib.host_f8(np.ones(3))
    args = [fake, fake, fake, 2, 4, 4, 4, op, dp, fake, 3, 1.0, 10, 0, fake, fake, fake, None]
    assert L.fteik3d_raytrace_batch(*args[:9], fake, 0, *args[11:]) == 0  # nothing to trace
    for i, bad in ((0, None), (1, None), (2, None), (3, 0), (4, 1), (11, 0.0), (12, 1)):
        a = list(args)
        a[i] = bad
        assert L.fteik3d_raytrace_batch(*a) == _lib.ERR_ARG, i


def test_graph_cache_query_without_device():
    """The kept-graph query takes no device work: nothing kept before any solve."""
    from fteikpy_amd import _lib

    L = _lib.lib()
    h, m = ctypes.c_longlong(-1), ctypes.c_longlong(-1)
    assert L.fteik3d_graph_cache(0, ctypes.byref(h), ctypes.byref(m)) == 0
    assert h.value >= 0 and m.value >= 0
    assert L.fteik3d_graph_cache(1, None, None) == 0
