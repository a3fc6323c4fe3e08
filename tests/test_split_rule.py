"""The split simulation's rule (ce_sim_split_slots): which list slots the one-lane kernel runs and
which the 2-lane kernel of the remainder runs.  The kernels' grids rely on these properties: the
split is a whole number of one-lane rounds (a multiple of 64 x cap slots, so the remainder's 2-lane
blocks start on a block boundary), and the remainder fits the 2-lane launch of cap blocks of 32
candidates."""
import pytest

from controllerengine_amd import _abi


def split(live, cap):
    return int(_abi.lib().ce_sim_split_slots(int(live), int(cap)))


def test_c3_block():
    # c3: 149,427 simulated candidates, 2,335 one-lane waves on 1,024 SIMDs: two full rounds and 287
    assert split(149427, 1024) == 2 * 1024 * 64


@pytest.mark.parametrize("live", [0, 1, 64, 65535, 65536, 131072, 131073 + 64 * 600, 2 * 131072, 10 ** 6])
def test_no_split_cases(live):
    cap = 1024
    w = (live + 63) // 64
    k, r = divmod(w, cap)
    s = split(live, cap)
    if k >= 2 and r > 0 and 2 * r <= cap:
        assert s < live
    else:
        assert s == live


@pytest.mark.parametrize("cap", [1, 4, 256, 1024])
def test_split_properties(cap):
    for live in range(0, 5 * 64 * cap + 1, max(1, 64 * cap // 37)):
        s = split(live, cap)
        assert 0 <= s <= live
        if s < live:
            assert s % (64 * cap) == 0 and s // (64 * cap) >= 2
            # the remainder: at most half a round, within the 2-lane grid (cap blocks of 32)
            assert live - s <= 32 * cap
        else:
            assert s == live


def test_no_devices_needed():
    # the rule is a host-callable function of the library: no HIP device is touched
    assert split(100, 0) == 100
