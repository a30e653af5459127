"""LDS-tiled Cartesian evaluation (split mode): bpmi_eval_grid and
bpmi_eval_grid_multi_tiled equal the per-pixel kernel (bpmi_eval_grid_multi
with 256-pixel blocks) bit for bit -- on grids whose tile windows fit in LDS,
on grids whose windows exceed it (global-memory path) and on ragged sizes --
and match the oracle's evaluate_image."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def root_image():
    from bp_project_amd.batched_resident_backend import BatchedResidentProgram
    from bp_project_amd.device import synchronize
    from bp_project_amd.gpu_backend import HipBaseline, build_ffbp_plan
    from bp_project_amd.scenes import polar_extent_for_grid, stripmap_scene
    s = stripmap_scene(pulses=1024, bins=2048, n_targets=8, grid_n=128, grid_spacing=0.5, seed=9)
    extent = polar_extent_for_grid(s.grid, s.pos.mean(axis=0), margin=2.0)
    plan = build_ffbp_plan(s.pos, s.fc, extent, side=1, leaf_size=16, range_spacing=0.5,
                           angular_oversample=2.5)
    backend = HipBaseline(interp="lagrange4", merge_interp="kaiser8")
    dev = backend.upload(s.data, s.pos, s.weight, s.phase, plan)
    prog = BatchedResidentProgram(backend, plan, dev, "split", s.r_start, s.data_dr, s.k)
    prog.execute()
    synchronize()
    return s, plan, prog


def _three_ways(prog, plan, k, grid, interp):
    from bp_project_amd import _lib
    from bp_project_amd.device import DeviceArray, stream_handle, synchronize
    from bp_project_amd.gpu_backend import geo_array
    from bp_project_amd.multipatch import EVAL_JOB
    x0, dx, nx, y0, dy, ny = grid
    vp = _lib.c_void_p
    geo = geo_array(plan)
    outs = [DeviceArray.empty((ny, nx)) for _ in range(3)]
    _lib.call("bpmi_eval_grid", vp(prog.output.ptr), geo.ctypes.data_as(vp), plan.nr, plan.nu,
              plan.nu, x0, dx, nx, y0, dy, ny, k, _lib.MODE_SPLIT, _lib.interp_code(interp),
              vp(outs[0].ptr), stream_handle())
    job = np.zeros(1, EVAL_JOB)
    job["img"], job["out"] = prog.output.ptr, outs[1].ptr
    job["geo"] = geo
    job["x0"], job["dx"], job["y0"], job["dy"], job["k"] = x0, dx, y0, dy, k
    job["nr"], job["nu"], job["ld"], job["nx"], job["ny"] = plan.nr, plan.nu, plan.nu, nx, ny
    jobs1 = DeviceArray.from_host(job.view(np.uint8))
    blocks = DeviceArray.from_host(np.array([(0, s) for s in range(0, nx * ny, 256)], np.int32))
    _lib.call("bpmi_eval_grid_multi", vp(jobs1.ptr), vp(blocks.ptr), blocks.shape[0],
              _lib.MODE_SPLIT, _lib.interp_code(interp), stream_handle())
    job["out"] = outs[2].ptr
    jobs2 = DeviceArray.from_host(job.view(np.uint8))
    t = ctypes.c_int()
    _lib.call("bpmi_eval_tiles", nx, ny, ctypes.byref(t))
    tiles = DeviceArray.from_host(np.array([(0, i) for i in range(t.value)], np.int32))
    _lib.call("bpmi_eval_grid_multi_tiled", vp(jobs2.ptr), vp(tiles.ptr), t.value,
              _lib.interp_code(interp), stream_handle())
    synchronize()
    return [o.get() for o in outs]


@pytest.mark.parametrize("interp", ["kaiser8", "lagrange4"])
@pytest.mark.parametrize("case", ["patch", "coarse_y", "ragged"])
def test_tiled_evaluation_equals_per_pixel(root_image, interp, case):
    s, plan, prog = root_image
    g = s.grid
    grid = {"patch": (g.x0, g.dx, g.nx, g.y0, g.dy, g.ny),
            # 20 m rows: a 4-row tile spans hundreds of angular cells, above the
            # LDS window cap, so every pixel takes the global-memory path
            "coarse_y": (g.x0, g.dx, 96, g.y0, 20.0, 3),
            "ragged": (g.x0 + 0.3, 0.37, 100, g.y0 + 1.1, 0.41, 7)}[case]
    tiled, per_pixel, multi_tiled = _three_ways(prog, plan, s.k, grid, interp)
    assert np.array_equal(tiled, per_pixel)
    assert np.array_equal(multi_tiled, per_pixel)
    assert np.abs(tiled).max() > 0


def test_tiled_evaluation_vs_oracle(oracle, root_image):
    s, plan, prog = root_image
    g = s.grid
    tiled, _, _ = _three_ways(prog, plan, s.k, (g.x0, g.dx, g.nx, g.y0, g.dy, g.ny), "kaiser8")
    img = prog.output.get()
    ref = oracle.evaluate_image(img.astype(np.complex128), plan.origin, plan.r0, plan.dr, plan.u0,
                                plan.du, g.points(), s.k, interp="kaiser8").reshape(g.shape)
    err = np.linalg.norm(tiled - ref) / np.linalg.norm(ref)
    print(f"tiled evaluation vs oracle evaluate_image: {err:.2e}")
    assert err <= 1e-5


@pytest.mark.parametrize("interp", ["kaiser8", "lagrange4"])
def test_points_on_the_node_origin_evaluate_to_zero(root_image, interp):
    """A point straight below or above the root origin has no polar angle
    (ground distance 0): its value is 0, in the point kernel and in a tiled
    grid whose middle pixel sits on the origin, and nothing is NaN."""
    from bp_project_amd.ffbp import evaluate_image
    s, plan, prog = root_image
    o = np.asarray(plan.origin, np.float64)
    img = prog.output.get()
    pts = np.array([[o[0], o[1], 0.0], [o[0], o[1], o[2]], [o[0] + 1.0, o[1], 0.0]])
    for mode in ("split", "fp64"):
        v = evaluate_image(img, plan.origin, plan.r0, plan.dr, plan.u0, plan.du, pts, s.k,
                           mode=mode, interp=interp)
        assert np.all(np.isfinite(v)), mode
        assert v[0] == 0 and v[1] == 0, mode
    d = 0.5
    tiled, per_pixel, multi = _three_ways(prog, plan, s.k, (o[0] - 2 * d, d, 5, o[1] - 2 * d, d, 5),
                                          interp)
    for out in (tiled, per_pixel, multi):
        assert np.all(np.isfinite(out))
        assert out[2, 2] == 0
