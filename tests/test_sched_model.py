"""CPU: the persistent 3D sweep schedule's dependency protocol, modelled with real threads
(tests/schedmodel/sched_model.cpp), which runs the kernel's own task decoding and
dependency slots (fteikpy_amd/csrc/schedule.hpp) with the kernel's memory orders.

The schedule is refused by the library: its first GPU run stalled.  These tests establish
what the CPU can: task indices decode into the grid, every (pass, tile, group) once per
set; progress indices stay inside the workspace; every wait ends; every brick read sees
the version the sequential order gives; ThreadSanitizer finds no access to the data that
the acquire/release chain leaves unordered -- on the stalled test's own geometry (21 x 19
x 27 cells, 3-plane bricks, 5 sources in groups of 2) and larger ones.  Dropping either
half of the waits is caught.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "schedmodel", "sched_model.cpp")


def _build(tmp, tsan):
    exe = str(tmp / ("sched_model_tsan" if tsan else "sched_model"))
    flags = ["-O1", "-g", "-fsanitize=thread"] if tsan else ["-O2"]
    subprocess.run(["g++", "-std=c++17", "-pthread", *flags, "-o", exe, SRC], check=True)
    return exe


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    return _build(tmp_path_factory.mktemp("schedmodel"), tsan=False)


@pytest.fixture(scope="module")
def model_tsan(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    try:
        return _build(tmp_path_factory.mktemp("schedmodel_tsan"), tsan=True)
    except subprocess.CalledProcessError:
        pytest.skip("ThreadSanitizer is not available for g++ here")


def _run(exe, *args, env=None):
    r = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, timeout=300,
                       env=env)
    return r.returncode, r.stdout + r.stderr


# (nodes z, x, y, brick height, sources, sources per workgroup, threads, sets, delay us)
CASES = [
    (22, 20, 28, 3, 5, 2, 8, 3, 0),     # the geometry of the first (stalled) GPU run
    (22, 20, 28, 3, 5, 2, 16, 2, 50),   # more workgroups than cores, completions reordered
    (22, 20, 28, 1, 5, 4, 8, 1, 0),     # one-plane bricks
    (41, 50, 45, 8, 9, 4, 12, 3, 20),   # face-byte tiles, partial bricks on every axis
    (161, 100, 100, 27, 4, 2, 16, 1, 0),  # the default height of the many-source workload
]


@pytest.mark.parametrize("case", CASES)
def test_protocol_holds(model, case):
    rc, out = _run(model, *case)
    assert rc == 0 and "OK" in out, out


@pytest.mark.parametrize("case", CASES[:2] + CASES[3:4])
def test_protocol_is_race_free_under_tsan(model_tsan, case):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    rc, out = _run(model_tsan, *case, env=env)
    assert rc == 0 and "ThreadSanitizer" not in out, out[-3000:]


@pytest.mark.parametrize("plant", [1, 2])
def test_dropped_waits_are_caught(model, plant):
    rc, out = _run(model, 22, 20, 28, 3, 5, 2, 8, 2, 50, plant)
    assert rc != 0 and "FAIL" in out, out
