"""Pulse-split direct BP of small grids (bpmi_bp_grid_ws): the C1 workload
(1024 pulses x 2048 bins onto 256 x 256 at 0.5 m) split over pulse chunks
matches the one-chunk launch to FP32 summation order, is bitwise repeatable,
honours BPMI_ACCUMULATE, falls back to the one-chunk launch bit for bit when
the workspace holds a single partial, and meets the oracle's bar."""
import ctypes

import numpy as np
import pytest

from bp_project_amd import _lib
from bp_project_amd.device import DeviceArray, synchronize

pytestmark = pytest.mark.gpu
_vp = _lib.c_void_p


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def c1():
    from bp_project_amd.scenes import workload_scene
    s = workload_scene("c1_direct_small")
    return s, DeviceArray.from_host(s.data), DeviceArray.from_host(s.pos)


def _grid(s, dev, p, out, interp, ws=None, ws_bytes=0, flags=0, pulses=None):
    g = s.grid
    pe = s.pulses if pulses is None else pulses
    _lib.call("bpmi_bp_grid_ws", _vp(dev.ptr), s.bins, _vp(p.ptr), None, None, 0, pe, g.x0, g.dx,
              g.nx, g.y0, g.dy, g.ny, 0.0, s.r_start, s.data_dr, s.k, _lib.MODE_SPLIT,
              _lib.interp_code(interp), flags, _vp(out.ptr), _vp(ws.ptr) if ws else None,
              ws_bytes, None)
    synchronize()
    return out.get()


def _workspace(s):
    n = ctypes.c_size_t()
    _lib.call("bpmi_bp_grid_workspace_bytes", s.grid.nx, s.grid.ny, s.pulses, ctypes.byref(n))
    assert n.value > 0, "a 256 x 256 grid should be split"
    return DeviceArray.empty((n.value // 8,)), n.value


@pytest.mark.parametrize("interp", ["linear", "lagrange4"])
def test_split_equals_one_chunk(c1, interp):
    s, dev, p = c1
    g = s.grid
    ws, nbytes = _workspace(s)
    print(f"{interp}: workspace {nbytes / 2**20:.1f} MiB = {nbytes // (g.nx * g.ny * 8)} partials",
          flush=True)
    one = _grid(s, dev, p, DeviceArray.empty(g.shape), interp)
    split = _grid(s, dev, p, DeviceArray.empty(g.shape), interp, ws, nbytes)
    err = rel(split, one)
    print(f"split vs one chunk: {err:.2e}")
    assert err <= 1e-6
    again = _grid(s, dev, p, DeviceArray.empty(g.shape), interp, ws, nbytes)
    assert np.array_equal(again, split)
    # ragged pulse count: 1000 pulses do not divide into equal chunks
    one_r = _grid(s, dev, p, DeviceArray.empty(g.shape), interp, pulses=1000)
    split_r = _grid(s, dev, p, DeviceArray.empty(g.shape), interp, ws, nbytes, pulses=1000)
    assert rel(split_r, one_r) <= 1e-6


def test_accumulate_and_small_workspace(c1):
    s, dev, p = c1
    g = s.grid
    ws, nbytes = _workspace(s)
    one = _grid(s, dev, p, DeviceArray.empty(g.shape), "linear")
    acc = DeviceArray.from_host(one)
    twice = _grid(s, dev, p, acc, "linear", ws, nbytes, flags=_lib.ACCUMULATE)
    assert rel(twice, 2 * one) <= 1e-6
    # room for one partial only: the one-chunk launch, bit for bit
    small = _grid(s, dev, p, DeviceArray.empty(g.shape), "linear", ws, g.nx * g.ny * 8 + 8)
    assert np.array_equal(small, one)


def test_split_vs_oracle(oracle, c1):
    from bp_project_amd.metrics import peak_phase_error, target_peaks
    s, dev, p = c1
    g = s.grid
    ws, nbytes = _workspace(s)
    img = _grid(s, dev, p, DeviceArray.empty(g.shape), "linear", ws, nbytes)
    ref = oracle.backproject_points(s.data, s.pos, g.points(), s.r_start, s.data_dr, s.k,
                                    s.weight, s.phase, interp="linear").reshape(g.shape)
    nrmse = rel(img, ref)
    ph = peak_phase_error(ref, img, target_peaks(np.abs(ref), g, s.targets))
    print(f"C1 pulse-split: NRMSE {nrmse:.3e}, max peak phase error {max(ph):.3e} rad")
    assert nrmse <= 1e-4
    assert max(ph) <= 0.01
