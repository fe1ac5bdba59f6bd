"""The real-time rig's schedule (bench.py, host only): pass phases and IR
update offsets of the channel groups keep every tick's load even."""
import collections
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def _phase_channels(k):
    ph = collections.Counter()
    for g in range(k):
        ph[bench.rt_phase(g)] += bench.rt_group_channels(g)
    return [ph[p] for p in range(bench.RT_PASS)]


def test_group_sizes_and_channels():
    big = bench.rt_large()
    assert big == bench.rt_balanced() * bench.RT_PASS
    assert all(bench.rt_group_channels(g) == bench.RT_GROUP for g in range(big))
    assert bench.rt_group_channels(big) == bench.RT_GROUP // 2
    assert bench.rt_channels(big) == big * bench.RT_GROUP
    assert bench.rt_channels(big + 12) == big * bench.RT_GROUP + 12 * bench.RT_GROUP // 2


def test_pass_phases_balanced():
    # every pass phase carries the same channels at the layout's balanced
    # points, and at most one smaller group more in between
    big = bench.rt_large()
    for k in (big, big + bench.RT_PASS, big + 2 * bench.RT_PASS):
        assert len(set(_phase_channels(k))) == 1, k
    for k in range(big, big + 2 * bench.RT_PASS):
        ch = _phase_channels(k)
        assert max(ch) - min(ch) <= bench.RT_GROUP // 2, (k, ch)
    # the phases with one group more are even (they pass on even ticks)
    ch = _phase_channels(big + 5)
    assert [p for p in range(bench.RT_PASS) if ch[p] > min(ch)] == [0, 2, 4, 6, 8]


def test_update_offsets_even():
    # at most two groups per offset of the 50-tick period while the groups
    # fit in two per offset, large pairs only on odd offsets
    big = bench.rt_large()
    for k in range(1, big + 28):
        per = collections.defaultdict(list)
        for g in range(k):
            per[bench.rt_update_offset(g)].append(g)
        assert max(len(v) for v in per.values()) <= 2, k
        for off, gs in per.items():
            if sum(1 for g in gs if g < big) == 2:
                assert off % 2 == 1, (k, off, gs)
