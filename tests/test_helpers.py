"""CPU: the thread-count knob (fteikpy/_helpers.py:4-27), with the reference's own pin
(tests/test_misc.py:4-7: set_num_threads(1) then get_num_threads() == 1)."""
import pytest

import fteikpy_amd
from fteikpy_amd import _helpers


@pytest.fixture
def restore_threads():
    old = fteikpy_amd.get_num_threads()
    yield
    fteikpy_amd.set_num_threads(old)


def test_one_thread_round_trip(restore_threads):
    fteikpy_amd.set_num_threads(1)
    assert fteikpy_amd.get_num_threads() == 1


def test_every_count_in_range_round_trips(restore_threads):
    for n in sorted({1, 2, _helpers._host_threads()}):
        fteikpy_amd.set_num_threads(n)
        assert fteikpy_amd.get_num_threads() == n


@pytest.mark.parametrize("bad, err", [(0, ValueError), (-3, ValueError), (10 ** 6, ValueError),
                                      (1.5, TypeError), (True, TypeError)])
def test_out_of_range_counts_are_refused(restore_threads, bad, err):
    fteikpy_amd.set_num_threads(1)
    with pytest.raises(err):
        fteikpy_amd.set_num_threads(bad)
    assert fteikpy_amd.get_num_threads() == 1
