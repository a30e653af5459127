"""Host side of the one-launch FFBP schedule (bp_project_amd/persistent.py):
the work list is topologically ordered, complete and sized like the
level-by-level launches; argument validation of bpmi_ffbp_persistent."""
import ctypes

import numpy as np
import pytest

from bp_project_amd import _lib
from bp_project_amd.persistent import (WORK_EVAL, WORK_LEAF, WORK_MERGE,
                                       check_order, work_items)
from bp_project_amd.plan import NI_HEIGHT, NI_LEFT, NI_NR, NI_NU, NI_RIGHT, flatten_plan


def _plan(pulses=512, leaf=16):
    from bp_project_amd.gpu_backend import build_ffbp_plan
    from bp_project_amd.scenes import polar_extent_for_grid, stripmap_scene
    s = stripmap_scene(pulses=pulses, bins=1024, n_targets=3, grid_n=48, grid_spacing=0.5, seed=3)
    extent = polar_extent_for_grid(s.grid, s.pos.mean(axis=0), margin=2.0)
    plan = build_ffbp_plan(s.pos, s.fc, extent, side=1, leaf_size=leaf, range_spacing=0.5,
                           angular_oversample=2.5)
    return plan, flatten_plan(plan)


def test_work_list_order_and_counts():
    plan, flat = _plan()
    ni = flat.node_int
    root = plan.root.index
    jobs = [(root, 48, 48)]
    items, node_tiles, job_blocks = work_items(ni, 5, jobs)
    assert check_order(items, ni, [root])
    # every tile of every node exactly once, evaluation blocks last
    for n in range(len(ni)):
        kind = WORK_LEAF if ni[n, NI_HEIGHT] == 0 else WORK_MERGE
        tiles = sorted(items[(items[:, 0] == kind) & (items[:, 1] == n), 2].tolist())
        assert tiles == list(range(node_tiles[n]))
    t = ctypes.c_int()
    _lib.call("bpmi_merge_tiles", int(ni[root, NI_NR]), int(ni[root, NI_NU]), ctypes.byref(t))
    assert node_tiles[root] == t.value
    assert job_blocks.tolist() == [1 * 12]  # 64 x 4-pixel tiles of a 48 x 48 grid
    ev = items[items[:, 0] == WORK_EVAL]
    assert ev[:, 2].tolist() == list(range(12))
    assert np.flatnonzero(items[:, 0] == WORK_EVAL).min() == len(items) - len(ev)
    # heights never decrease along the list
    h = [ni[i[1], NI_HEIGHT] if i[0] != WORK_EVAL else ni[root, NI_HEIGHT] for i in items]
    assert all(a <= b for a, b in zip(h, h[1:]))


def test_check_order_rejects_a_parent_before_its_child():
    plan, flat = _plan()
    ni = flat.node_int
    items, _, _ = work_items(ni, 6)
    parent = int(np.flatnonzero(ni[:, NI_HEIGHT] == 1)[0])
    child = int(ni[parent, NI_LEFT])
    i_parent = int(np.flatnonzero((items[:, 0] == WORK_MERGE) & (items[:, 1] == parent))[0])
    i_child = int(np.flatnonzero((items[:, 0] == WORK_LEAF) & (items[:, 1] == child))[-1])
    bad = items.copy()
    bad[[i_parent, i_child]] = bad[[i_child, i_parent]]
    assert not check_order(bad, ni, [])
    assert ni[parent, NI_RIGHT] >= 0


def test_priority_moves_a_forest_s_small_tree_first():
    """Two trees in one table: with the smaller tree's nodes prioritised, its
    evaluation tiles come right after its root's height."""
    _, big = _plan(pulses=512)
    _, small = _plan(pulses=128)
    ni_b, ni_s = big.node_int.copy(), small.node_int.copy()
    base = len(ni_b)
    for col in (NI_LEFT, NI_RIGHT):
        ni_s[:, col] = np.where(ni_s[:, col] >= 0, ni_s[:, col] + base, -1)
    ni = np.concatenate([ni_b, ni_s])
    root_b = int(np.flatnonzero(ni_b[:, NI_HEIGHT] == ni_b[:, NI_HEIGHT].max())[0])
    root_s = base + int(np.flatnonzero(ni_s[:, NI_HEIGHT] == ni_s[:, NI_HEIGHT].max())[0])
    prio = np.r_[np.full(len(ni_b), 1), np.full(len(ni_s), 0)]
    items, _, _ = work_items(ni, 6, [(root_b, 10, 10), (root_s, 10, 10)], priority=prio)
    assert check_order(items, ni, [root_b, root_s])
    first_eval = items[np.flatnonzero(items[:, 0] == WORK_EVAL)[0]]
    assert first_eval[1] == 1  # the small tree's job
    leaves = items[items[:, 0] == WORK_LEAF, 1]
    assert leaves[0] >= base  # small tree's leaves first


def test_persistent_entry_point_validation():
    nbytes = ctypes.c_size_t()
    _lib.call("bpmi_persistent_workspace_bytes", 10, 2, ctypes.byref(nbytes))
    assert nbytes.value >= 512 + 4 * (2 + 10 + 2) + 8 * 3
    sc = _lib.PersistentSchedule()
    with pytest.raises(_lib.BpmiError, match="persistent merges"):
        _lib.call("bpmi_ffbp_persistent", ctypes.byref(sc), 0, 1, 0, None)
    sc.nitems = 4
    with pytest.raises(ValueError, match="null table"):
        _lib.call("bpmi_ffbp_persistent", ctypes.byref(sc), 0, 0, 0, None)


def test_work_list_without_leaves():
    plan, flat = _plan()
    ni = flat.node_int
    items, node_tiles, _ = work_items(ni, 6, leaves=False)
    assert not (items[:, 0] == WORK_LEAF).any()
    assert (node_tiles[ni[:, NI_HEIGHT] == 0] == 0).all()
    assert check_order(items, ni, [], leaves=False)
    assert not check_order(items, ni, [], leaves=True)


def test_persistent_without_leaves_is_switched_off():
    from bp_project_amd.persistent import PersistentFFBP
    plan, flat = _plan()
    with pytest.raises(NotImplementedError, match="switched off"):
        PersistentFFBP(flat.node_int, 0, 0, 0, 0, 1.0, 6, leaves=False)
    sc = _lib.PersistentSchedule()
    with pytest.raises(_lib.BpmiError, match="switched off"):
        _lib.call("bpmi_ffbp_persistent", ctypes.byref(sc), -1, 2, 0, None)
