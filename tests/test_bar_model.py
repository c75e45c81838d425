"""CPU: the persistent schedule's grid-wide barrier (tree of counters, generation poll) in
a thread model (tests/schedmodel/bar_model.cpp) that compiles the layout and counter
choice of fteikpy_amd/csrc/schedule.hpp: every barrier ends, every write before it is seen
after it, every counter is back at zero after each one, for fan-in edge cases and the
workgroup counts the library launches; ThreadSanitizer finds no unordered access; the line
layout fits the reserved lines for every workgroup count up to the maximum; a workgroup
that never arrives makes every other one leave with the give-up flag set (the launch ends
instead of hanging)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "schedmodel", "bar_model.cpp")


def _build(tmp, tsan):
    exe = str(tmp / ("bar_model_tsan" if tsan else "bar_model"))
    flags = ["-fsanitize=thread", "-O1", "-g"] if tsan else ["-O2"]
    subprocess.run(["g++", "-std=c++17", "-Wall", *flags, "-o", exe, SRC, "-pthread"],
                   check=True)
    return exe


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    return _build(tmp_path_factory.mktemp("bar"), tsan=False)


def test_layout_fits_every_workgroup_count(exe):
    r = subprocess.run([exe, "layout"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.returncode


@pytest.mark.parametrize("blocks", [1, 2, 15, 16, 17, 255, 256, 257, 1000])
def test_barrier_rounds(exe, blocks):
    r = subprocess.run([exe, str(blocks), "20"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.returncode


def test_barrier_under_tsan(tmp_path_factory):
    try:
        tsan = _build(tmp_path_factory.mktemp("bar_tsan"), tsan=True)
    except subprocess.CalledProcessError:
        pytest.skip("ThreadSanitizer not available")
    r = subprocess.run([tsan, "33", "10"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr[-2000:]


@pytest.mark.parametrize("blocks", [2, 17, 300])
def test_missing_workgroup_gives_up(exe, blocks):
    r = subprocess.run([exe, "giveup", str(blocks)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.returncode
