"""Autotuner logic on the CPU: the coordinate search over knob groups (one
choice per group or one per stage), the JSON cache (hits, forced re-runs,
stale and corrupted entries), the single-knob tuner and the launch-shape
validation of the level-launch programs.  Measurements are injected, so no
GPU is touched."""
import json

import pytest

from bp_project_amd import _lib
from bp_project_amd import autotune as at
from bp_project_amd.batched_resident_backend import check_launch_shapes


@pytest.fixture
def cache_file(tmp_path, monkeypatch):
    p = tmp_path / "tune.json"
    monkeypatch.setenv("BPMI_TUNE_CACHE", str(p))
    return p


def cost_model(calls):
    """Stage times of a made-up program: the leaf stage is cheapest at
    (5, 16) and costs more when the merges use shape B (a shared resource),
    merge stage 1 prefers shape A, stage 2 shape B."""
    def measure(cfg):
        calls.append(json.loads(json.dumps(cfg)))
        tc, pb = cfg["leaf"]
        leaves = abs(tc - 5) * 2.0 + abs(pb - 16) / 16.0 + 1.0
        m1 = {"A": 1.0, "B": 3.0, "C": 2.0}
        m2 = {"A": 4.0, "B": 2.5, "C": 3.0}
        s1, s2 = cfg["merges"]
        return {"leaves": leaves, "merges1": m1[s1], "merges2": m2[s2], "eval": 0.5}
    return measure


def groups():
    return [at.Group("leaf", [[tc, pb] for tc in (4, 5, 6) for pb in (16, 32, 64)], ["leaves"]),
            at.Group("merges", ["A", "B", "C"], ["merges1", "merges2"], per_stage=True)]


def test_coordinate_search_picks_group_and_stage_minima():
    calls = []
    base = {"leaf": [6, 32], "merges": ["A", "A"]}
    cfg, report = at.coordinate_search(groups(), cost_model(calls), base)
    assert cfg == {"leaf": [5, 16], "merges": ["A", "B"]}
    # 9 leaf candidates, then 3 merge shapes (each applied to every height)
    assert len(calls) == 12
    # the merge group was measured with the leaf group's winner in place
    assert all(c["leaf"] == [5, 16] for c in calls[9:])
    assert [c["merges"] for c in calls[9:]] == [["A", "A"], ["B", "B"], ["C", "C"]]
    assert report["merges"]["times"][json.dumps("B")] == [3.0, 2.5]
    assert report["leaf"]["choice"] == [5, 16]
    assert report["merges"]["stages"] == ["merges1", "merges2"]


def test_coordinate_search_rejects_bad_base_and_missing_stages():
    with pytest.raises(ValueError, match="no valid value"):
        at.coordinate_search(groups(), cost_model([]), {"leaf": [7, 32], "merges": ["A", "A"]})
    with pytest.raises(ValueError, match="no valid value"):
        at.coordinate_search(groups(), cost_model([]), {"leaf": [5, 32], "merges": ["A"]})

    def short(cfg):
        return {"leaves": 1.0}
    with pytest.raises(RuntimeError, match="lacks stages"):
        at.coordinate_search(groups(), short, {"leaf": [5, 32], "merges": ["A", "A"]})
    with pytest.raises(ValueError):
        at.Group("x", [], ["s"])
    with pytest.raises(ValueError):
        at.Group("x", [1], [])


def test_tune_config_caches_and_revalidates(cache_file):
    calls = []
    base = {"leaf": [6, 32], "merges": ["A", "A"]}
    cfg, report, cached = at.tune_config("k1", groups(), cost_model(calls), base)
    assert not cached and cfg == {"leaf": [5, 16], "merges": ["A", "B"]}
    n = len(calls)
    cfg2, report2, cached2 = at.tune_config("k1", groups(), cost_model(calls), base)
    assert cached2 and cfg2 == cfg and report2 == json.loads(json.dumps(report))
    assert len(calls) == n  # a cache hit measures nothing
    # force re-measures
    at.tune_config("k1", groups(), cost_model(calls), base, force=True)
    assert len(calls) == 2 * n
    # a cached value that is no longer a candidate is stale: re-measured
    g = [at.Group("leaf", [[tc, pb] for tc in (4, 6) for pb in (16, 32)], ["leaves"]),
         groups()[1]]
    cfg3, _, cached3 = at.tune_config("k1", g, cost_model(calls), base)
    assert not cached3 and cfg3["leaf"] in ([4, 16], [6, 16])
    # a cached per-stage list of the wrong length is stale too
    entry = json.loads(cache_file.read_text())
    entry["k1"]["config"]["merges"] = ["A"]
    cache_file.write_text(json.dumps(entry))
    m = len(calls)
    _, _, cached4 = at.tune_config("k1", groups(), cost_model(calls), base)
    assert not cached4 and len(calls) > m
    # keys are independent
    _, _, cached5 = at.tune_config("k2", groups(), cost_model(calls), base)
    assert not cached5
    assert set(json.loads(cache_file.read_text())) == {"k1", "k2"}


def test_corrupted_or_unwritable_cache_is_ignored(cache_file, tmp_path, monkeypatch):
    cache_file.write_text("{not json")
    calls = []
    cfg, _, cached = at.tune_config("k", groups(), cost_model(calls),
                                    {"leaf": [4, 64], "merges": ["C", "C"]})
    assert not cached and cfg["leaf"] == [5, 16]
    cache_file.write_text("[1, 2, 3]")  # valid JSON, wrong shape
    _, _, cached = at.tune_config("k", groups(), cost_model(calls),
                                  {"leaf": [4, 64], "merges": ["C", "C"]})
    assert not cached
    # a cache path that cannot be written: the search still returns
    monkeypatch.setenv("BPMI_TUNE_CACHE", str(tmp_path / "file_as_dir" / "x" / "tune.json"))
    (tmp_path / "file_as_dir").write_text("")
    cfg, _, cached = at.tune_config("k", groups(), cost_model([]),
                                    {"leaf": [4, 64], "merges": ["C", "C"]})
    assert not cached and cfg["merges"] == ["A", "B"]


def test_single_knob_tune_with_injected_timer(cache_file):
    seen = []

    def timer(run, c, warmup, repeats):
        seen.append(c)
        return {4: 3.0, 5: 1.0, 6: 2.0}[c]
    best, times = at.tune("w", [4, 5, 6], None, timer=timer)
    assert best == 5 and times == {"4": 3.0, "5": 1.0, "6": 2.0} and seen == [4, 5, 6]
    best, _ = at.tune("w", [4, 5, 6], None, timer=timer)
    assert best == 5 and len(seen) == 3  # cached
    best, _ = at.tune("w", [4, 6], None, timer=timer)  # cached winner not offered: re-run
    assert best == 6 and len(seen) == 5


def test_keys_and_size_classes():
    assert at.size_class(1) == 0 and at.size_class(4) == 1 and at.size_class(4 ** 10) == 10
    assert at.size_class(0) == 0
    k = at.make_key("multipatch_levels", "split", "lagrange4+kaiser8", 4 ** 12,
                    arch="gfx950:sramecc+:xnack-", extra="h7")
    assert k == "gfx950/multipatch_levels/split/lagrange4+kaiser8/s12/h7"
    assert at.stage_keys(["leaves", "merges", "merges", "eval"]) == [
        "leaves", "merges1", "merges2", "eval"]
    assert at.stage_keys(["leaves", "merges", "eval"]) == ["leaves", "merges", "eval"]


def test_launch_shape_validation():
    pb, shapes = check_launch_shapes("split", None, None, 3)
    assert pb == _lib.DEFAULT_PULSE_BLOCK == 32
    assert shapes == [_lib.DEFAULT_MERGE_SHAPE] * 3 == [(4, 2176)] * 3
    assert check_launch_shapes("split", 64, (5, 1664), 2) == (64, [(5, 1664)] * 2)
    assert check_launch_shapes("split", 16, [[7, 1088], [5, 1920]], 2)[1] == [(7, 1088),
                                                                               (5, 1920)]
    with pytest.raises(ValueError, match="pulse_block must be one of"):
        check_launch_shapes("split", 48, None, 2)
    with pytest.raises(ValueError, match="merge shape must be one of"):
        check_launch_shapes("split", None, (3, 2176), 2)
    with pytest.raises(ValueError, match="one shape per merge height"):
        check_launch_shapes("split", None, [[4, 2176]] * 3, 2)
    with pytest.raises(ValueError, match="split mode only"):
        check_launch_shapes("mixed", 16, None, 2)
    assert set(_lib.PULSE_BLOCKS) == {16, 32, 64}
    assert len(set(_lib.MERGE_SHAPES)) == len(_lib.MERGE_SHAPES) == 4


def test_c_abi_shape_lists_match_the_python_constants():
    """The library validates launch shapes before touching a device; its
    messages list exactly the shapes the Python side offers the tuner."""
    import ctypes
    lib = _lib.load()
    one = ctypes.c_void_p(1)
    # lagrange4 merges need no Kaiser table, so the shape check is the first
    # to fail
    st = lib.bpmi_ffbp_merges_tiled_ex(one, one, one, one, one, 1, 1.0,
                                       _lib.INTERP_CODES["lagrange4"], 3, 2176, None)
    assert st == _lib.ERR_INVALID
    msg = _lib.last_error()
    listed = [tuple(int(v) for v in m) for m in
              __import__("re").findall(r"\((\d+), (\d+)\)", msg.split("one of", 1)[1])]
    assert listed == list(_lib.MERGE_SHAPES)
    st = lib.bpmi_ffbp_leaves_multi_ex(one, one, one, one, one, one, 1, _lib.MODE_SPLIT,
                                       _lib.INTERP_CODES["lagrange4"], 6, 48, None)
    assert st == _lib.ERR_INVALID
    assert _lib.last_error() == "pulse_block must be 16, 32 or 64"
    assert sorted(_lib.PULSE_BLOCKS) == [16, 32, 64]


def test_bench_autotune_summary_handles_both_report_shapes():
    """The bench's tuner record accepts the coordinate-search report and the
    single-knob report of a persistent program (one whole-launch stage)."""
    import bench
    calls = []
    cfg, report = at.coordinate_search(groups(), cost_model(calls),
                                       {"leaf": [6, 32], "merges": ["A", "A"]})
    out = bench.autotune_summary(cfg, report, False, {"leaf": [6, 32], "merges": "A"},
                                 per_stage=("merges",))
    assert out["default_vs_tuned"]["merges"] == {"default_ms": 5.0, "tuned_ms": 3.5}
    assert out["default_vs_tuned"]["leaf"]["tuned_ms"] <= out["default_vs_tuned"]["leaf"]["default_ms"]
    persistent = {"leaf": {"stages": ["batch"], "choice": [5, None],
                           "times": {json.dumps([4, None]): [3.0], json.dumps([5, None]): [2.0]}}}
    out = bench.autotune_summary({"leaf": [5, None]}, persistent, False, {"leaf": [6, 32]})
    assert out["candidates_ms"]["leaf"]["times"] == {"[4, null]": [3.0], "[5, null]": [2.0]}
    assert out["default_vs_tuned"] == {}
