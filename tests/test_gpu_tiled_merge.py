"""Split-mode LDS-tiled merge (bpmi_ffbp_merges_tiled, and
bpmi_ffbp_merges_tiled_win reading a bpmi_ffbp_merge_windows table) on
hand-built node tables (also with a window table whose bounds belong to
other tiles, which the kernel clips to each child): bitwise equal to the
untiled split merge (bpmi_ffbp_merges) and
within 1e-4 of the FP64-mode merge, both when each child's footprint fits
the LDS tile (coarse child u sampling) and when it does not and the kernel
gathers from global memory (child u sampling 8x finer than the parent's).
Part of the parent lies outside the children's valid stencil range, so
zero-filled pixels are covered too, and so is a tile containing a pixel on
a child's origin."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p


def _tables(child_du):
    from bp_project_amd.plan import NI_HEIGHT, NI_LEFT, NI_NR, NI_NU, NI_RIGHT, NI_SIDE
    parent = dict(o=(0.0, 0.0, 3000.0), r0=5500.0, dr=0.5, nr=45, u0=-0.02, du=1e-3, nu=45)
    nu_c = int(np.ceil(0.07 / child_du)) + 1
    kids = [dict(o=(0.0, s * 20.0, 3000.0), r0=5501.0, dr=0.5, nr=220, u0=-0.035, du=child_du,
                 nu=nu_c) for s in (-1.0, 1.0)]
    nodes = [parent] + kids
    node_int = np.zeros((3, 8), np.int32)
    node_geo = np.zeros((3, 7), np.float64)
    offsets = np.zeros(3, np.int64)
    cursor = 0
    for i, n in enumerate(nodes):
        node_int[i, NI_NR], node_int[i, NI_NU], node_int[i, NI_SIDE] = n["nr"], n["nu"], 1
        node_int[i, NI_LEFT] = node_int[i, NI_RIGHT] = -1
        node_geo[i] = (*n["o"], n["r0"], n["dr"], n["u0"], n["du"])
        offsets[i] = cursor
        cursor += -(-n["nr"] * n["nu"] // 64) * 64
    node_int[0, NI_LEFT], node_int[0, NI_RIGHT], node_int[0, NI_HEIGHT] = 1, 2, 1
    return nodes, node_int, node_geo, offsets, cursor


def _window_table(t_int, t_geo, blocks, interp, node_int, table):
    """bpmi_ffbp_merge_windows of a block table, checked on the host: every
    window lies inside its child, a stageable one is at least one stencil
    wide, and the reciprocal spacings are the children's to FP64 rounding."""
    from bp_project_amd.batched_resident_backend import merge_window_table
    from bp_project_amd.plan import NI_LEFT, NI_NR, NI_NU, NI_RIGHT
    windows = merge_window_table(t_int.ptr, t_geo.ptr, blocks, interp)
    w = windows.get()
    geo = t_geo.get()
    taps = 8 if interp == "kaiser8" else 4
    for (node, _), e in zip(table, w):
        if node < 0:
            assert not e.any()
            continue
        kids = (node_int[node, NI_LEFT], node_int[node, NI_RIGHT])
        inv = e[8:].view(np.float64).reshape(2, 2)
        for child, (idr, idu) in zip(kids, inv):
            assert abs(idr * geo[child, 4] - 1) < 1e-15 and abs(idu * geo[child, 6] - 1) < 1e-15
        for child, (r0, c0, h, wd) in zip(kids, e[:8].reshape(2, 4)):
            assert (h == 0) == (wd == 0)
            if h:
                assert h >= taps and wd >= taps and r0 >= 0 and c0 >= 0
                assert r0 + h <= node_int[child, NI_NR] and c0 + wd <= node_int[child, NI_NU]
    return windows


@pytest.mark.parametrize("child_du", [2e-3, 1.25e-4])
@pytest.mark.parametrize("interp", ["kaiser8", "lagrange4"])
def test_tiled_merge_matches_untiled(child_du, interp):
    from bp_project_amd import _lib
    from bp_project_amd.batched_resident_backend import merge_block_table
    from bp_project_amd.device import DeviceArray, synchronize
    from bp_project_amd.gpu_backend import ensure_kaiser_table
    ensure_kaiser_table()
    nodes, node_int, node_geo, offsets, size = _tables(child_du)
    rng = np.random.default_rng(7)
    arena_host = np.zeros(size, np.complex64)
    for i in (1, 2):
        n = nodes[i]["nr"] * nodes[i]["nu"]
        arena_host[offsets[i]:offsets[i] + n] = (rng.standard_normal(n) +
                                                 1j * rng.standard_normal(n)).astype(np.complex64)
    t_int, t_geo = DeviceArray.from_host(node_int), DeviceArray.from_host(node_geo)
    t_off = DeviceArray.from_host(offsets)
    k = 4 * np.pi * 9.6e9 / 299792458.0
    p0 = nodes[0]["nr"] * nodes[0]["nu"]
    results = {}
    for name in ("tiled", "tiled_win", "tiled_moved", "untiled", "fp64"):
        arena = DeviceArray.from_host(arena_host)
        mode = "fp64" if name == "fp64" else "split"
        kind = "split" if name.startswith("tiled") else "x"
        table = np.asarray(merge_block_table(node_int, [0], kind), np.int32)
        blocks = DeviceArray.from_host(table)
        if name == "tiled":
            _lib.call("bpmi_ffbp_merges_tiled", _vp(t_int.ptr), _vp(t_geo.ptr), _vp(t_off.ptr),
                      _vp(arena.ptr), _vp(blocks.ptr), len(table), k, _lib.interp_code(interp),
                      None)
        elif name.startswith("tiled_"):
            windows = _window_table(t_int, t_geo, blocks, interp, node_int, table)
            if name == "tiled_moved":
                # bounds of other tiles, a huge window and one past the
                # image's corner: clipped to each child, only slower
                host = windows.get()
                bounds = host[:, :8].copy()
                host[:, :8] = np.roll(bounds, 1, axis=0)
                host[0, :4] = (0, 0, 1 << 20, 1 << 20)
                host[-1, 4:8] = (-50, 10_000, 30, 30)
                windows = DeviceArray.from_host(host)
            _lib.call("bpmi_ffbp_merges_tiled_win", _vp(t_int.ptr), _vp(t_geo.ptr),
                      _vp(t_off.ptr), _vp(arena.ptr), _vp(blocks.ptr), len(table), k,
                      _lib.interp_code(interp), *_lib.DEFAULT_MERGE_SHAPE, _vp(windows.ptr),
                      None)
        else:
            _lib.call("bpmi_ffbp_merges", _vp(t_int.ptr), _vp(t_geo.ptr), _vp(t_off.ptr),
                      _vp(arena.ptr), _vp(blocks.ptr), len(table), k, _lib.mode_code(mode),
                      _lib.interp_code(interp), None)
        synchronize()
        results[name] = arena.get()[:p0]
    a, b, c = results["tiled"], results["untiled"], results["fp64"]
    zero = np.count_nonzero(a == 0)
    print(f"child du {child_du} {interp}: {zero} of {p0} parent pixels outside the children")
    assert 0 < zero < p0 // 2
    assert np.array_equal(a, b)
    assert np.array_equal(results["tiled_win"], a)
    assert np.array_equal(results["tiled_moved"], a)
    assert np.linalg.norm(a - c) <= 1e-4 * np.linalg.norm(c)


@pytest.mark.parametrize("interp", ["kaiser8", "lagrange4"])
def test_tile_with_a_pixel_on_a_child_origin(interp):
    """Parent origin (0, 0, 0), children at (0, 10, 0): parent pixel (3, 2)
    (r = 9.25 + 3 * 0.25 = 10, u = 0.5 + 2 * 0.25 = 1) lies on both children's
    origin, where the tiled merge's geometry produces NaN (rsqrt64(0)) and the
    pixel must read nothing: equal to the untiled merge bit for bit, the
    origin pixel 0, nothing NaN."""
    from bp_project_amd import _lib
    from bp_project_amd.batched_resident_backend import merge_block_table
    from bp_project_amd.device import DeviceArray, synchronize
    from bp_project_amd.gpu_backend import ensure_kaiser_table
    from bp_project_amd.plan import NI_HEIGHT, NI_LEFT, NI_NR, NI_NU, NI_RIGHT, NI_SIDE
    ensure_kaiser_table()
    geo = [(0.0, 0.0, 0.0, 9.25, 0.25, 0.5, 0.25), (0.0, 10.0, 0.0, 0.5, 0.25, -0.9, 0.03),
           (0.0, 10.0, 0.0, 0.5, 0.25, -0.9, 0.03)]
    shapes = [(8, 3), (96, 64), (96, 64)]
    node_int = np.zeros((3, 8), np.int32)
    offsets = np.zeros(3, np.int64)
    cursor = 0
    for i, (nr, nu) in enumerate(shapes):
        node_int[i, NI_NR], node_int[i, NI_NU], node_int[i, NI_SIDE] = nr, nu, 1
        node_int[i, NI_LEFT] = node_int[i, NI_RIGHT] = -1
        offsets[i] = cursor
        cursor += -(-nr * nu // 64) * 64
    node_int[0, NI_LEFT], node_int[0, NI_RIGHT], node_int[0, NI_HEIGHT] = 1, 2, 1
    node_geo = np.array(geo, np.float64)
    rng = np.random.default_rng(3)
    arena_host = np.zeros(cursor, np.complex64)
    for i in (1, 2):
        n = 96 * 64
        arena_host[offsets[i]:offsets[i] + n] = (rng.standard_normal(n) +
                                                 1j * rng.standard_normal(n)).astype(np.complex64)
    t_int, t_geo = DeviceArray.from_host(node_int), DeviceArray.from_host(node_geo)
    t_off = DeviceArray.from_host(offsets)
    k = 4 * np.pi * 9.6e9 / 299792458.0
    out = {}
    for name in ("tiled", "tiled_win", "untiled"):
        arena = DeviceArray.from_host(arena_host)
        kind = "split" if name.startswith("tiled") else "x"
        table = np.asarray(merge_block_table(node_int, [0], kind), np.int32)
        blocks = DeviceArray.from_host(table)
        if name == "tiled":
            _lib.call("bpmi_ffbp_merges_tiled", _vp(t_int.ptr), _vp(t_geo.ptr), _vp(t_off.ptr),
                      _vp(arena.ptr), _vp(blocks.ptr), len(table), k, _lib.interp_code(interp),
                      None)
        elif name == "tiled_win":
            windows = _window_table(t_int, t_geo, blocks, interp, node_int, table)
            _lib.call("bpmi_ffbp_merges_tiled_win", _vp(t_int.ptr), _vp(t_geo.ptr),
                      _vp(t_off.ptr), _vp(arena.ptr), _vp(blocks.ptr), len(table), k,
                      _lib.interp_code(interp), *_lib.DEFAULT_MERGE_SHAPE, _vp(windows.ptr),
                      None)
        else:
            _lib.call("bpmi_ffbp_merges", _vp(t_int.ptr), _vp(t_geo.ptr), _vp(t_off.ptr),
                      _vp(arena.ptr), _vp(blocks.ptr), len(table), k, _lib.MODE_SPLIT,
                      _lib.interp_code(interp), None)
        synchronize()
        out[name] = arena.get()[:24].reshape(8, 3)
    assert np.all(np.isfinite(out["tiled"]))
    assert out["tiled"][3, 2] == 0
    assert np.count_nonzero(out["tiled"]) >= 6
    assert np.array_equal(out["tiled"], out["untiled"])
    assert np.array_equal(out["tiled_win"], out["tiled"])
