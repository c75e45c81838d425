"""CPU: the overlapped 3D sweep schedule's brick list (fteikpy_amd/csrc/schedule.hpp),
dumped by tests/schedmodel/overlap_dump.cpp, against a model of what each brick touches.

A brick of pass p on memory tile T reads the nodes and change bytes of T and of T's 7
upwind tiles in pass p, and writes only T's.  The reference sweeps each pass over the whole
grid before the next (fteikpy/_fteik/_fteik3d.py:217-391); the pass-by-pass brick order
(one launch per anti-diagonal) reproduces it.  The model gives every write a version and
records, for every brick, the version of each tile it reads.  Replaying the overlapped
launches with every brick of a launch reading before any of them writes -- the worst
interleaving a launch allows -- must record exactly the versions of the pass-by-pass
order, and so must the opposite extreme (bricks of a launch one after another, in reverse
order).  A list that moves one brick a launch too early is caught.
"""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "schedmodel", "overlap_dump.cpp")


@pytest.fixture(scope="module")
def dump(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("overlap") / "overlap_dump")
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-o", exe, SRC], check=True)

    def run(nba, nbb, nbc):
        r = subprocess.run([exe, str(nba), str(nbb), str(nbc)], capture_output=True, text=True,
                           check=True, timeout=120)
        lines = r.stdout.split("\n")
        nl, nt = map(int, lines[0].split())
        launches = [[tuple(map(int, b.split(","))) for b in ln.split()]
                    for ln in lines[1:1 + nl]]
        assert sum(map(len, launches)) == nt
        return launches
    return run


def _signs(p):
    # pass order of the reference: x flips every pass, y every 2, z every 4
    return (-1 if p & 4 else 1, -1 if p & 1 else 1, -1 if p & 2 else 1)


def _tile(p, b, dims):
    return tuple(v if s > 0 else n - 1 - v for v, s, n in zip(b, _signs(p), dims))


def _reads(p, t, dims):
    s = _signs(p)
    out = [t]
    for q in range(1, 8):
        u = tuple(t[a] - (s[a] if q >> a & 1 else 0) for a in range(3))
        if all(0 <= u[a] < dims[a] for a in range(3)):
            out.append(u)
    return out


def _sequential(dims):
    """Bricks pass by pass, anti-diagonal by anti-diagonal (oriented coordinates)."""
    nba, nbb, nbc = dims
    order = []
    for p in range(8):
        bricks = [(a, b, c) for a in range(nba) for b in range(nbb) for c in range(nbc)]
        bricks.sort(key=lambda x: sum(x))
        order += [(p, *b) for b in bricks]
    return order


def _replay(launches, dims, together):
    """Versions each brick reads; together: a launch's bricks all read before any writes."""
    ver, seen = {}, {}
    for launch in launches:
        if together:
            for (p, *b) in launch:
                t = _tile(p, b, dims)
                seen[(p, t)] = tuple(ver.get(u) for u in _reads(p, t, dims))
            for (p, *b) in launch:
                ver[_tile(p, b, dims)] = (p, _tile(p, b, dims))
        else:
            for (p, *b) in reversed(launch):
                t = _tile(p, b, dims)
                seen[(p, t)] = tuple(ver.get(u) for u in _reads(p, t, dims))
                ver[t] = (p, t)
    return seen


DIMS = [(1, 1, 1), (1, 3, 2), (2, 2, 2), (3, 5, 4), (4, 3, 6), (6, 9, 7), (2, 11, 3)]


@pytest.mark.parametrize("dims", DIMS)
def test_overlapped_reads_match_pass_by_pass(dump, dims):
    launches = dump(*dims)
    bricks = [b for launch in launches for b in launch]
    nba, nbb, nbc = dims
    assert sorted(bricks) == sorted((p, a, b, c) for p in range(8) for a in range(nba)
                                    for b in range(nbb) for c in range(nbc))
    ref = _replay([[b] for b in _sequential(dims)], dims, together=False)
    assert _replay(launches, dims, together=True) == ref
    assert _replay(launches, dims, together=False) == ref
    assert len(launches) <= 8 * (nba + nbb + nbc - 2)


def test_a_brick_moved_earlier_is_caught(dump):
    dims = (3, 5, 4)
    launches = dump(*dims)
    ref = _replay([[b] for b in _sequential(dims)], dims, together=False)
    caught = 0
    for l in range(1, len(launches)):
        moved = [list(x) for x in launches]
        brick = moved[l].pop(0)
        moved[l - 1].append(brick)
        caught += (_replay(moved, dims, together=True) != ref or
                   _replay(moved, dims, together=False) != ref)
    assert caught == len(launches) - 1


def test_launch_counts(dump):
    """Launches per sweep set: 508 instead of 688 for the 160 x 320 x 320 cell grid in
    27-plane bricks (6 x 41 x 41 tiles), 368 instead of 608 for 200^3 cells in 8-plane
    bricks (26^3 tiles); a single tile still needs its 8 passes in order."""
    assert len(dump(6, 41, 41)) == 508
    assert len(dump(26, 26, 26)) == 368
    assert len(dump(1, 1, 1)) == 8
    counts = np.array([len(x) for x in dump(6, 41, 41)])
    assert counts.sum() == 8 * 6 * 41 * 41
