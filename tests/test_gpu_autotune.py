"""Launch shapes the autotuner chooses from compute what the default shapes
compute: every merge shape (occupancy hint, LDS window cap) bit for bit,
every leaf pulse block depth and tile width to FP32 rounding; a tuned batch
equals the per-patch programs with the same leaf shapes bit for bit, and
the tuner's record names every candidate it timed."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def scenes():
    from bp_project_amd.multipatch import make_patches, patch_mix
    return make_patches(patch_mix(count=6, seed=5, scale=0.125))


def run(prog):
    from bp_project_amd.device import synchronize
    prog.execute()
    synchronize()
    return [o.get() for o in prog.outputs]


def rel(a, b):
    return max(float(np.linalg.norm(x - y) / np.linalg.norm(y)) for x, y in zip(a, b))


@pytest.mark.parametrize("merge_interp", ["kaiser8", "lagrange4"])
def test_every_launch_shape_matches_the_default(scenes, merge_interp):
    from bp_project_amd import _lib
    from bp_project_amd.multipatch import MultiPatchProgram
    prog = MultiPatchProgram(scenes, mode="split", merge_interp=merge_interp, tile_cols_log2=6)
    ref = run(prog)
    heights = len(prog.merge_tables)
    assert prog.launch_config() == {"leaf": [6, 32], "merge_shapes": [[4, 2176]] * heights}
    for shape in _lib.MERGE_SHAPES[1:]:
        prog.set_launch_config({"leaf": [6, 32], "merge_shapes": [list(shape)] * heights})
        out = run(prog)
        assert all(np.array_equal(a, b) for a, b in zip(out, ref)), shape
    mixed = [list(_lib.MERGE_SHAPES[h % len(_lib.MERGE_SHAPES)]) for h in range(heights)]
    prog.set_launch_config({"leaf": [6, 32], "merge_shapes": mixed})
    assert all(np.array_equal(a, b) for a, b in zip(run(prog), ref))
    for tc in (4, 5, 6):
        for pb in _lib.PULSE_BLOCKS:
            prog.set_launch_config({"leaf": [tc, pb], "merge_shapes": mixed})
            d = rel(run(prog), ref)
            print(f"{merge_interp} leaf tile 2^{tc} pulse block {pb}: {d:.2e}", flush=True)
            assert d <= 1e-5, (tc, pb, d)
    prog.close()


def test_tuned_batch_equals_sequential_and_records_candidates(scenes, tmp_path, monkeypatch):
    from bp_project_amd import _lib
    from bp_project_amd.multipatch import MultiPatchProgram, sequential_reference
    monkeypatch.setenv("BPMI_TUNE_CACHE", str(tmp_path / "tune.json"))
    prog = MultiPatchProgram(scenes, mode="split", merge_interp="kaiser8", cull=True)
    ref = run(prog)
    config, report, cached = prog.autotune(force=True)
    assert not cached
    heights = len(prog.merge_tables)
    tc, pb = config["leaf"]
    assert tc in (4, 5, 6) and pb in _lib.PULSE_BLOCKS
    assert len(config["merge_shapes"]) == heights
    assert all(tuple(s) in _lib.MERGE_SHAPES for s in config["merge_shapes"])
    assert prog.launch_config() == config
    assert len(report["leaf"]["times"]) == 9
    assert len(report["merge_shapes"]["times"]) == len(_lib.MERGE_SHAPES)
    assert all(len(v) == heights and min(v) > 0
               for v in report["merge_shapes"]["times"].values())
    out = run(prog)
    assert rel(out, ref) <= 1e-5
    seq, _, _, progs = sequential_reference(scenes, mode="split", merge_interp="kaiser8",
                                            tile_cols_log2=tc, cull=True, pulse_block=pb)
    assert all(np.array_equal(a, b.get()) for a, b in zip(out, seq))
    for _, p_, _, _ in progs:
        p_.close()
    # the second tune is a cache hit with the same choice
    again, _, cached = prog.autotune()
    assert cached and again == config
    assert json.loads((tmp_path / "tune.json").read_text())
    # under graph replay the tuned shapes hold
    assert prog.capture_graph()
    prog.replay()
    from bp_project_amd.device import synchronize
    synchronize()
    assert all(np.array_equal(a, o.get()) for a, o in zip(out, prog.outputs))
    prog.close()


def test_single_plan_launch_shapes_and_tuner(tmp_path, monkeypatch):
    """BatchedResidentProgram: shapes set in place on the built schedule
    (merge shapes bitwise, pulse blocks to FP32 rounding), and its tuner."""
    from bp_project_amd import _lib
    from bp_project_amd.batched_resident_backend import BatchedResidentProgram
    from bp_project_amd.device import synchronize
    from bp_project_amd.gpu_backend import HipBaseline, build_ffbp_plan
    from bp_project_amd.scenes import polar_extent_for_grid, stripmap_scene
    monkeypatch.setenv("BPMI_TUNE_CACHE", str(tmp_path / "tune.json"))
    s = stripmap_scene(pulses=1024, bins=2048, n_targets=8, grid_n=128, grid_spacing=0.5, seed=7)
    extent = polar_extent_for_grid(s.grid, s.pos.mean(axis=0), margin=2.0)
    plan = build_ffbp_plan(s.pos, s.fc, extent, side=1, leaf_size=16, range_spacing=0.5,
                           angular_oversample=2.5)
    backend = HipBaseline(interp="lagrange4", merge_interp="kaiser8")
    dev = backend.upload(s.data, s.pos, s.weight, s.phase, plan)
    prog = BatchedResidentProgram(backend, plan, dev, "split", s.r_start, s.data_dr, s.k)
    prog.execute()
    synchronize()
    ref = prog.output.get()
    levels = len(prog.merge_shapes)
    for shape in _lib.MERGE_SHAPES[1:]:
        prog.set_launch_shapes(merge_shapes=shape)
        prog.execute()
        synchronize()
        assert np.array_equal(prog.output.get(), ref), shape
    built = BatchedResidentProgram(backend, plan, dev, "split", s.r_start, s.data_dr, s.k,
                                   pulse_block=64,
                                   merge_shapes=[list(_lib.MERGE_SHAPES[-1])] * levels)
    built.execute()
    synchronize()
    got = built.output.get()
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
    assert [e[1].cname for e in built._schedule[:2]] == ["bpmi_ffbp_leaves_multi_ex",
                                                         "bpmi_ffbp_merges_tiled_win"]
    built.close()
    config, report, cached = prog.autotune(force=True)
    assert not cached and config["pulse_block"] in _lib.PULSE_BLOCKS
    assert len(config["merge_shapes"]) == levels
    prog.execute()
    synchronize()
    assert np.linalg.norm(prog.output.get() - ref) <= 1e-5 * np.linalg.norm(ref)
    with pytest.raises(ValueError, match="merge shape must be one of"):
        prog.set_launch_shapes(merge_shapes=(3, 2176))
    prog.close()


def test_shape_arguments_rejected_at_the_boundary():
    import ctypes
    from bp_project_amd import _lib
    lib = _lib.load()
    one = ctypes.c_void_p(1)
    st = lib.bpmi_ffbp_merges_tiled_ex(one, one, one, one, one, 1, 1.0, 0, 3, 2176, None)
    assert st == _lib.ERR_INVALID and "merge shape" in _lib.last_error()
    st = lib.bpmi_ffbp_leaves_multi_ex(one, one, one, one, one, one, 1, 4, 0, 6, 48, None)
    assert st == _lib.ERR_INVALID and "pulse_block must be 16, 32 or 64" in _lib.last_error()
