"""One-launch (persistent) FFBP on the GPU: bitwise equal to the level-by-level
launches for a single plan and for many uneven patches, under graph replay
too; a dependency wait that runs out is reported and the launch still ends."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _single(merge_interp, persistent):
    from bp_project_amd.batched_resident_backend import BatchedResidentProgram
    from bp_project_amd.gpu_backend import HipBaseline, build_ffbp_plan
    from bp_project_amd.scenes import polar_extent_for_grid, stripmap_scene
    s = stripmap_scene(pulses=1024, bins=2048, n_targets=8, grid_n=128, grid_spacing=0.5, seed=5)
    extent = polar_extent_for_grid(s.grid, s.pos.mean(axis=0), margin=2.0)
    plan = build_ffbp_plan(s.pos, s.fc, extent, side=1, leaf_size=16, range_spacing=0.5,
                           angular_oversample=2.5)
    backend = HipBaseline(interp="lagrange4", merge_interp=merge_interp)
    dev = backend.upload(s.data, s.pos, s.weight, s.phase, plan)
    prog = BatchedResidentProgram(backend, plan, dev, "split", s.r_start, s.data_dr, s.k,
                                  leaf_tile_cols_log2=5, persistent=persistent)
    return s, plan, prog


MERGES_ONLY_OFF = pytest.mark.skip(
    reason="persistent list without leaves switched off: its one GPU run did not finish "
           "(test timeout at its first launch), cause not found; the path raises")


@pytest.mark.parametrize("merge_interp,persistent", [
    ("kaiser8", True), ("lagrange4", True),
    pytest.param("kaiser8", "merges", marks=MERGES_ONLY_OFF)])
def test_single_plan_persistent_equals_level_launches(oracle, merge_interp, persistent):
    from bp_project_amd.device import synchronize
    s, plan, ref_prog = _single(merge_interp, False)
    _, _, prog = _single(merge_interp, persistent)
    print(f"{merge_interp}: {prog.persistent.nitems} work items, launches {prog.launch_count}")
    kept = 0 if persistent is True else 1
    assert prog.launch_count == ref_prog.launch_count - len(ref_prog.flat.levels) + 1 + kept
    ref_prog.execute()
    synchronize()
    ref = ref_prog.output.get()
    for rep in range(3):
        prog.execute()
        synchronize()
        assert not prog.persistent.aborted()
        assert np.array_equal(prog.output.get(), ref), f"replay {rep} differs"
    ff, _ = oracle.form_ffbp(s.data, s.pos, s.r_start, s.data_dr, s.fc, plan.extent, side=1,
                             leaf_size=16, range_spacing=0.5, angular_oversample=2.5,
                             weight=s.weight, phase=s.phase, merge_interp=merge_interp)
    err = np.linalg.norm(ref - ff.image) / np.linalg.norm(ff.image)
    print(f"vs FP64 oracle FFBP: {err:.2e}")
    assert err <= 1e-4
    prog.close()
    ref_prog.close()


def _patches():
    from bp_project_amd.multipatch import make_patches, patch_mix
    return make_patches(patch_mix(count=10, scale=0.25))


@pytest.mark.parametrize("persistent", [True, pytest.param("merges", marks=MERGES_ONLY_OFF)])
def test_multipatch_persistent_equals_level_launches_and_graph(persistent):
    from bp_project_amd.device import synchronize
    from bp_project_amd.multipatch import MultiPatchProgram
    scenes = _patches()
    ref_prog = MultiPatchProgram(scenes, mode="split", tile_cols_log2=5)
    prog = MultiPatchProgram(scenes, mode="split", tile_cols_log2=5, persistent=persistent)
    assert prog.launch_count == (1 if persistent is True else 2)
    ref_prog.execute()
    synchronize()
    ref = [o.get() for o in ref_prog.outputs]
    prog.execute()
    synchronize()
    got = [o.get() for o in prog.outputs]
    assert all(np.array_equal(a, b) for a, b in zip(got, ref))
    t = prog.completion_ms()
    print("patch completion ms:", np.round(t, 3))
    assert np.all(t > 0) and np.all(np.isfinite(t))
    prog.capture_graph()
    for rep in range(3):
        for o in prog.outputs:
            o.fill_zero()
        ref_prog.execute()  # eager work between replays
        prog.replay()
        synchronize()
        assert all(np.array_equal(o.get(), r) for o, r in zip(prog.outputs, ref)), rep
        assert not prog._persistent.aborted()
    prog.close()
    ref_prog.close()


def test_persistent_wait_timeout_is_reported_and_the_launch_ends():
    from bp_project_amd.device import synchronize
    from bp_project_amd.multipatch import MultiPatchProgram
    from bp_project_amd.plan import NI_HEIGHT
    scenes = _patches()[:3]
    prog = MultiPatchProgram(scenes, mode="split", tile_cols_log2=5, persistent=True)
    prog.execute()
    synchronize()
    good = [o.get() for o in prog.outputs]
    pf = prog._persistent
    leaf = int(np.flatnonzero(prog.node_int[:, NI_HEIGHT] == 0)[0])
    tiles = pf.node_tiles_host.copy()
    tiles[leaf] += 1  # a completion count the leaf never reaches
    pf.node_tiles_dev.set(tiles)
    pf.schedule.max_wait_s = 0.5
    t = time.perf_counter()
    prog.execute()
    synchronize()
    elapsed = time.perf_counter() - t
    assert pf.aborted()
    # the bound is in wall-clock time: the first wait that cannot complete
    # gives up after 0.5 s at the device's reported clock rate
    print(f"aborted launch took {elapsed:.2f} s for a 0.5 s wait bound")
    assert 0.4 <= elapsed <= 5.0
    with pytest.raises(RuntimeError, match="aborted"):
        prog.completion_ms()
    pf.node_tiles_dev.set(pf.node_tiles_host)
    pf.schedule.max_wait_s = 10.0
    prog.execute()
    synchronize()
    assert not pf.aborted()
    assert all(np.array_equal(o.get(), g) for o, g in zip(prog.outputs, good))
    prog.close()
