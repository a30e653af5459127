"""The persistent FFBP kernel's scheduling protocol, simulated on the host
(tests/persistent_protocol_sim.cpp: the kernel's chunk claims, pending
completion counts, flushes and bounded dependency waits, with threads as
resident workgroup slots) over real work lists: the merges-only list of the
case whose one GPU run did not finish (1024 pulses, leaf 16, 695 items),
the same plan with its leaves, and a 10-patch forest with Cartesian
evaluation jobs; for the workgroup counts of a launch (occupancy 4 and 5 on
256 CUs), far fewer resident slots than workgroups, and random item
durations.  Every run must execute every item once, never abort and never
start an item before its dependencies completed.  Two planted mistakes must
end in the abort: a work list out of dependency order, and no flush of the
pending count before a workgroup waits."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from bp_project_amd.plan import NI_LEFT, NI_RIGHT, flatten_plan
from bp_project_amd.persistent import work_items

HERE = Path(__file__).resolve().parent


@pytest.fixture(scope="module")
def sim(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("sim") / "persistent_protocol_sim"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", str(HERE / "persistent_protocol_sim.cpp"),
                    "-o", str(exe)], check=True)
    return exe


def _plan(pulses=1024, grid_n=128, seed=5):
    from bp_project_amd.gpu_backend import build_ffbp_plan
    from bp_project_amd.scenes import polar_extent_for_grid, stripmap_scene
    s = stripmap_scene(pulses=pulses, bins=64, n_targets=0, grid_n=grid_n, grid_spacing=0.5,
                       seed=seed, snr_db=None)
    extent = polar_extent_for_grid(s.grid, s.pos.mean(axis=0), margin=2.0)
    return build_ffbp_plan(s.pos, s.fc, extent, side=1, leaf_size=16, range_spacing=0.5,
                           angular_oversample=2.5), s.grid


def _text(node_int, items, node_tiles, jobs=(), job_blocks=()):
    lines = [f"{len(node_int)} {len(items)} {len(jobs)}"]
    lines += [f"{int(n[NI_LEFT])} {int(n[NI_RIGHT])} {int(t)}" for n, t in zip(node_int, node_tiles)]
    lines += [f"{int(k)} {int(i)}" for k, i, *_ in items]
    lines += [f"{int(r)} {int(b)}" for (r, _, _), b in zip(jobs, job_blocks)]
    return "\n".join(lines) + "\n"


def _run(sim, text, G, R, seed=0, wait_ms=2000, delay_us=30, *extra):
    p = subprocess.run([str(sim), str(G), str(R), str(seed), str(wait_ms), str(delay_us), *extra],
                       input=text, capture_output=True, text=True, timeout=300)
    return p.returncode, p.stdout.strip()


@pytest.fixture(scope="module")
def single():
    plan, _ = _plan()
    return flatten_plan(plan).node_int


@pytest.fixture(scope="module")
def forest():
    node_int, jobs, base = [], [], 0
    for i in range(10):
        plan, grid = _plan(pulses=256 * (1 + i % 4), grid_n=64 + 32 * (i % 3), seed=100 + i)
        ni = flatten_plan(plan).node_int.copy()
        for col in (NI_LEFT, NI_RIGHT):
            ni[:, col] = np.where(ni[:, col] >= 0, ni[:, col] + base, -1)
        node_int.append(ni)
        jobs.append((base + plan.root.index, grid.nx, grid.ny))
        base += len(ni)
    return np.concatenate(node_int), jobs


@pytest.mark.parametrize("G,R", [(1024, 1024), (1280, 1280), (1280, 256), (1024, 16)])
def test_merges_only_list_of_the_unfinished_run_completes(sim, single, G, R):
    items, tiles, _ = work_items(single, 5, leaves=False)
    assert len(items) == 695
    for seed in range(2):
        rc, out = _run(sim, _text(single, items, tiles), G, R, seed)
        assert rc == 0, out


@pytest.mark.parametrize("G,R", [(1024, 1024), (1280, 64)])
def test_list_with_leaves_completes(sim, single, G, R):
    items, tiles, _ = work_items(single, 5, leaves=True)
    rc, out = _run(sim, _text(single, items, tiles), G, R, 1)
    assert rc == 0, out


def test_forest_with_evaluation_jobs_completes(sim, forest):
    node_int, jobs = forest
    for leaves in (True, False):
        items, tiles, blocks = work_items(node_int, 5, jobs=jobs, leaves=leaves)
        rc, out = _run(sim, _text(node_int, items, tiles, jobs, blocks), 1024, 128, 2)
        assert rc == 0, out


def test_planted_mistakes_end_in_the_abort(sim, single):
    items, tiles, _ = work_items(single, 5, leaves=True)
    # dependency order broken: the list reversed, one resident workgroup
    rc, out = _run(sim, _text(single, items[::-1], tiles), 4, 1, 0, 300, 0)
    assert rc == 1 and "aborted 1" in out, out
    # no flush before a wait: a workgroup that holds the last tile count of
    # a child in its pending count waits for it forever
    rc, out = _run(sim, _text(single, items, tiles), 1, 1, 0, 300, 0, "noflush")
    assert rc == 1 and "aborted 1" in out, out
