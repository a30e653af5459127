"""Block-table construction for the tiled merges (host logic, no GPU)."""
import numpy as np

from bp_project_amd.batched_resident_backend import xcd_order


def test_xcd_order_is_a_padded_permutation():
    for n in (1, 7, 8, 9, 100, 1027):
        entries = [(i // 5, i % 5) for i in range(n)]
        out = xcd_order(entries)
        assert len(out) % 8 == 0 and len(out) - n < 8
        real = [e for e in out if e[0] >= 0]
        assert sorted(real) == sorted(entries)          # every tile exactly once
        assert all(e == (-1, 0) for e in out if e[0] < 0)


def test_xcd_order_gives_each_die_a_contiguous_run():
    entries = list(range(61))
    out = xcd_order([(e, 0) for e in entries])
    for x in range(8):
        run = [e[0] for e in out[x::8] if e[0] >= 0]    # workgroups x, x + 8, ..
        assert run == list(range(run[0], run[0] + len(run)))
    firsts = [out[x][0] for x in range(8) if out[x][0] >= 0]
    assert firsts == sorted(firsts)
    assert np.diff(firsts).min() >= 7


def test_group_by_work():
    """Latency groups: indices ordered by work (ties by index), cut at
    rounded fractions of the count, every index once, no empty group."""
    from bp_project_amd.multipatch import group_by_work
    work = [50, 10, 30, 10, 90, 70, 20, 60]
    g = group_by_work(work, (0.5, 0.75))
    assert g == [[1, 3, 6, 2], [0, 7], [5, 4]]
    assert group_by_work(work, ()) == [[1, 3, 6, 2, 0, 7, 5, 4]]
    assert group_by_work(work, (0.0, 1.0, 0.01)) == [[1, 3, 6, 2, 0, 7, 5, 4]]
    assert group_by_work([5], (0.5,)) == [[0]]
    assert [len(x) for x in group_by_work(list(range(64)), (0.5, 0.75))] == [32, 16, 16]
