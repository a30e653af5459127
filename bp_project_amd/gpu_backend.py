"""HIP backend with the call signatures of fast_bp/gpu_backend.py.

``HipBaseline`` (also exported as ``CudaBaseline`` so callers of the
reference switch by changing the import) uploads inputs once and runs direct
BP and FFBP through the libbpmi C ABI on the current HIP stream.

Modes: the reference's ``MODES`` = ("fp64", "fp32", "mixed") plus
``"dfloat"`` (double-float range and phase) and ``"split"`` (the LDS-tiled
kernel: FP64 per-tile geometry reference, FP32 residual, FP32 interpolation
and accumulation).  Every mode shares one FP64 plan (``build_ffbp_plan``).

Differences from the reference that are deliberate (not CUDA limitations):
* node images and accumulators are ``DeviceArray`` objects (``cp.ndarray``
  needs CUDA), with the same shape/dtype/``data.ptr``/``get()`` surface;
* the kernels take the FP64 position/weight/phase tables and round per mode
  on the device; the FP32 copies are still uploaded so ``DeviceInputs`` has
  every reference field.
"""
from __future__ import annotations

import platform
import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _lib
from .device import (DeviceArray, Event, POOL, current_stream, device_info, elapsed_ms,
                     stream_handle)
from .plan import FFBPPlan, NodePlan, build_ffbp_plan, iter_nodes  # noqa: F401 (re-export)

MODES = ("fp64", "fp32", "mixed")
EXTRA_MODES = ("dfloat", "split")
ALL_MODES = MODES + EXTRA_MODES

_c = _lib
_vp = _lib.c_void_p


def check_mode(mode):
    """The reference's mode error (fast_bp/gpu_backend.py:442-443) over the
    accepted tuple: MODES plus this build's dfloat and split."""
    if mode not in ALL_MODES:
        raise ValueError(f"mode must be one of {ALL_MODES}")



@dataclass
class DeviceInputs:
    """Device copies of one acquisition (fast_bp/gpu_backend.py:105-118)."""
    data: DeviceArray
    pos64: DeviceArray
    pos32: DeviceArray
    weight64: DeviceArray
    weight32: DeviceArray
    phase64: DeviceArray
    phase32: DeviceArray
    origins64: DeviceArray
    origins32: DeviceArray
    h2d_event_ms: float
    h2d_wall_ms: float
    bytes_uploaded: int


def geo_array(node):
    """Host FP64 geometry row (ox, oy, oz, r0, dr, u0, du) of a node/plan."""
    return np.ascontiguousarray([node.origin[0], node.origin[1], node.origin[2], node.r0,
                                 node.dr, node.u0, node.du], dtype=np.float64)


def _gp(geo):
    return geo.ctypes.data_as(_vp)


def origin_reference(position, a, b, r_min):
    """Split-mode reference choice for a polar node: the node origin when the
    node's antennas stay within 2% of the nearest slant range (short
    sub-apertures, e.g. FFBP leaves), else each tile's centre."""
    pos = np.asarray(position[a:b], dtype=np.float64)
    origin = pos.mean(axis=0)
    spread = float(np.sqrt(((pos - origin) ** 2).sum(axis=1)).max()) if len(pos) else 0.0
    return int(spread <= 0.02 * max(float(r_min), 1e-9))


SPLIT_WINDOW_SLOTS = 32   # LDS range window of the split kernel (samples per pulse)


def leaf_tile_width(plan, position, data_dr):
    """log2 tile width (16..64 u-columns) of the split-mode leaf kernel for a
    plan: the least padded u-columns over all leaves whose tile rows, widened
    by the leaf sub-aperture, still fit the LDS range window (a pulse that does
    not fit takes the exact-geometry path, so this is a speed choice only).
    Eager, fused and batched FFBP all use it, so they stay bitwise equal."""
    leaves = [n for n in iter_nodes(plan.root) if n.is_leaf]
    position = np.asarray(position, dtype=np.float64)
    reach = 0.0
    for n in leaves:
        p = position[n.a:n.b]
        reach = max(reach, float(np.sqrt(((p - p.mean(axis=0)) ** 2).sum(axis=1)).max()))
    best, best_cols = 6, None
    for tc in (4, 5, 6):
        rows = 1024 >> tc
        slots = (rows * leaves[0].dr + 2 * reach) / float(data_dr) + 4
        if slots > SPLIT_WINDOW_SLOTS:
            continue
        cols = sum(-(-n.nu // (1 << tc)) * (1 << tc) for n in leaves)
        if best_cols is None or cols < best_cols:
            best, best_cols = tc, cols
    return best


class HipBaseline:
    """Launch the libbpmi kernels (counterpart of fast_bp/gpu_backend.py:227-561)."""

    kernel_names = (
        "bpmi_bp_polar", "bpmi_bp_polar_ex", "bpmi_bp_grid", "bpmi_bp_points", "bpmi_regrid",
        "bpmi_merge2", "bpmi_add_inplace", "bpmi_cast_c128_c64", "bpmi_crop",
        "bpmi_eval_points", "bpmi_eval_grid", "bpmi_ffbp_leaves", "bpmi_ffbp_merges",
        "bpmi_target_contributions", "bpmi_fill_range_phase_lut", "bpmi_bp_polar_lut",
    )

    def __init__(self, block_size=128, interp="lagrange4", merge_interp="lagrange4"):
        if block_size <= 0:
            raise ValueError("block_size must be positive")
        self.block_size = int(block_size)
        start = time.perf_counter()
        lib = _lib.load()
        if _device_count() == 0:
            raise RuntimeError("no HIP device visible: the HIP backend has no CPU fallback")
        self.kernels = {name: getattr(lib, name) for name in self.kernel_names}
        self.compile_seconds = time.perf_counter() - start
        self.source_path = Path(__file__).resolve().parent / "csrc"
        self.library_path = _lib.LIB_PATH
        self.interp = interp
        self.merge_interp = merge_interp
        if "kaiser8" in (interp, merge_interp):
            ensure_kaiser_table()
        self.peak_pool_used_bytes = POOL.peak

    def _track_memory(self):
        self.peak_pool_used_bytes = max(self.peak_pool_used_bytes, POOL.peak)

    # -- transfers -------------------------------------------------------------
    def upload(self, data, position, weight, phase, plan):
        data = np.ascontiguousarray(data) if not data.flags.c_contiguous else data
        if data.dtype != np.complex64 or data.ndim != 2:
            raise TypeError("range-compressed data must be C-contiguous complex64 [pulse, range]")
        position64 = np.ascontiguousarray(position, dtype=np.float64)
        weight64 = np.ascontiguousarray(weight, dtype=np.float64)
        phase64 = np.ascontiguousarray(phase, dtype=np.float64)
        if position64.shape != (data.shape[0], 3):
            raise ValueError("position shape does not match the pulse count")
        if weight64.shape != (data.shape[0],) or phase64.shape != (data.shape[0],):
            raise ValueError("weight and phase must have one value per pulse")
        host_start = time.perf_counter()
        start, end = Event(), Event()
        start.record()
        origins = plan.origins if plan is not None else np.zeros((1, 3))
        device = DeviceInputs(
            data=DeviceArray.from_host(data),
            pos64=DeviceArray.from_host(position64),
            pos32=DeviceArray.from_host(position64.astype(np.float32)),
            weight64=DeviceArray.from_host(weight64),
            weight32=DeviceArray.from_host(weight64.astype(np.float32)),
            phase64=DeviceArray.from_host(phase64),
            phase32=DeviceArray.from_host(phase64.astype(np.float32)),
            origins64=DeviceArray.from_host(np.ascontiguousarray(origins, dtype=np.float64)),
            origins32=DeviceArray.from_host(np.ascontiguousarray(origins, dtype=np.float32)),
            h2d_event_ms=0.0, h2d_wall_ms=0.0, bytes_uploaded=0)
        end.record()
        end.synchronize()
        device.h2d_event_ms = elapsed_ms(start, end)
        device.h2d_wall_ms = 1000.0 * (time.perf_counter() - host_start)
        device.bytes_uploaded = int(sum(getattr(device, f).nbytes for f in (
            "data", "pos64", "pos32", "weight64", "weight32", "phase64", "phase32", "origins64",
            "origins32")))
        device._host_position = position64
        self._track_memory()
        return device

    @staticmethod
    def download(array):
        start = time.perf_counter()
        result = array.get()
        return result, 1000.0 * (time.perf_counter() - start)

    # -- building blocks ---------------------------------------------------------
    def leaf(self, node, device, mode, range_start, data_dr, k, out=None, stream=None,
             tile_cols_log2=-1):
        """Leaf polar BP of one node into ``out`` (new array if None)."""
        out = DeviceArray.empty((node.nr, node.nu)) if out is None else out
        geo = geo_array(node)
        d = device
        if mode == "split":
            oref = origin_reference(d._host_position, node.a, node.b, node.r0)
            _lib.call("bpmi_bp_polar_ex", _vp(d.data.ptr), d.data.shape[1], _vp(d.pos64.ptr),
                      _vp(d.weight64.ptr), _vp(d.phase64.ptr), node.a, node.b, _gp(geo), node.nr,
                      node.nu, node.side, range_start, data_dr, k, _lib.MODE_SPLIT,
                      _lib.interp_code(self.interp), 0, _vp(out.ptr), stream_handle(stream),
                      oref, tile_cols_log2)
        else:
            _lib.call("bpmi_bp_polar", _vp(d.data.ptr), d.data.shape[1], _vp(d.pos64.ptr),
                      _vp(d.weight64.ptr), _vp(d.phase64.ptr), node.a, node.b, _gp(geo), node.nr,
                      node.nu, node.side, range_start, data_dr, k, _lib.mode_code(mode),
                      _lib.interp_code(self.interp), 0, _vp(out.ptr), stream_handle(stream))
        return out

    def regrid(self, image, source, destination, mode, k, out=None, accumulate=False, stream=None):
        out = DeviceArray.empty((destination.nr, destination.nu)) if out is None else out
        sg, dg = geo_array(source), geo_array(destination)
        _lib.call("bpmi_regrid", _vp(image.ptr), _gp(sg), source.nr, source.nu, _gp(dg),
                  destination.nr, destination.nu, destination.side, k, _lib.mode_code(mode),
                  _lib.interp_code(self.merge_interp), _lib.ACCUMULATE if accumulate else 0,
                  _vp(out.ptr), stream_handle(stream))
        return out

    def merge2(self, left_image, left, right_image, right, destination, mode, k, out=None,
               stream=None):
        out = DeviceArray.empty((destination.nr, destination.nu)) if out is None else out
        lg, rg, dg = geo_array(left), geo_array(right), geo_array(destination)
        _lib.call("bpmi_merge2", _vp(left_image.ptr), _gp(lg), left.nr, left.nu,
                  _vp(right_image.ptr), _gp(rg), right.nr, right.nu, _gp(dg), destination.nr,
                  destination.nu, destination.side, k, _lib.mode_code(mode),
                  _lib.interp_code(self.merge_interp), _vp(out.ptr), stream_handle(stream))
        return out

    @staticmethod
    def add_inplace(left, right, stream=None):
        _lib.call("bpmi_add_inplace", _vp(left.ptr), _vp(right.ptr), left.size,
                  stream_handle(stream))

    @staticmethod
    def crop(root, plan, out=None, stream=None):
        out = DeviceArray.empty(plan.shape) if out is None else out
        _lib.call("bpmi_crop", _vp(root.ptr), plan.root.nu, plan.crop_r[0], plan.crop_u[0],
                  plan.nr, plan.nu, _vp(out.ptr), stream_handle(stream))
        return out

    # -- eager schedules -----------------------------------------------------------
    def _execute_node(self, node, device, mode, range_start, data_dr, k, tc=-1):
        if node.is_leaf:
            return self.leaf(node, device, mode, range_start, data_dr, k, tile_cols_log2=tc)
        left_child = self._execute_node(node.left, device, mode, range_start, data_dr, k, tc)
        left = self.regrid(left_child, node.left, node, mode, k)
        del left_child
        right_child = self._execute_node(node.right, device, mode, range_start, data_dr, k, tc)
        right = self.regrid(right_child, node.right, node, mode, k)
        del right_child
        self.add_inplace(left, right)
        del right
        self._track_memory()
        return left

    def run_ffbp(self, plan, device, mode, range_start, data_dr, k):
        """Eager depth-first FFBP (fast_bp/gpu_backend.py:441-469)."""
        check_mode(mode)
        wall_start = time.perf_counter()
        start, end = Event(), Event()
        start.record()
        tc = leaf_tile_width(plan, device._host_position, data_dr) if mode == "split" else -1
        root = self._execute_node(plan.root, device, mode, range_start, data_dr, k, tc)
        output = self.crop(root, plan)
        del root
        end.record()
        end.synchronize()
        event_ms = elapsed_ms(start, end)
        wall_ms = 1000.0 * (time.perf_counter() - wall_start)
        self._track_memory()
        return output, event_ms, wall_ms

    def run_direct(self, plan, device, mode, range_start, data_dr, k, chunk_pulses=64):
        """Direct BP on the cropped root grid, accumulated over pulse chunks
        (fast_bp/gpu_backend.py:471-555)."""
        check_mode(mode)
        if chunk_pulses <= 0:
            raise ValueError("chunk_pulses must be positive")
        count = plan.nr * plan.nu
        if mode == "fp64":
            accumulator = DeviceArray.zeros(count, np.complex128)
            output = DeviceArray.empty(count, np.complex64)
        else:
            accumulator = DeviceArray.zeros(count, np.complex64)
            output = accumulator
        pulses = int(device.data.shape[0])
        wall_start = time.perf_counter()
        start, end = Event(), Event()
        start.record()
        self.enqueue_direct(plan, device, mode, range_start, data_dr, k, accumulator,
                            chunk_pulses=chunk_pulses)
        if mode == "fp64":
            _lib.call("bpmi_cast_c128_c64", _vp(accumulator.ptr), _vp(output.ptr), count,
                      stream_handle())
        end.record()
        end.synchronize()
        event_ms = elapsed_ms(start, end)
        wall_ms = 1000.0 * (time.perf_counter() - wall_start)
        del accumulator
        self._track_memory()
        del pulses
        return output.reshape(plan.shape), event_ms, wall_ms

    def enqueue_direct(self, plan, device, mode, range_start, data_dr, k, accumulator,
                       chunk_pulses=64, stream=None):
        """Enqueue the chunked direct BP into a zeroed accumulator."""
        d = device
        pulses = int(d.data.shape[0])
        geo = geo_array(plan)
        flags = _lib.ACCUMULATE | (_lib.OUT_C128 if mode == "fp64" else 0)
        oref = 0
        if mode == "split":
            oref = origin_reference(d._host_position, 0, pulses, plan.r0)
        for begin in range(0, pulses, int(chunk_pulses)):
            finish = min(pulses, begin + int(chunk_pulses))
            _lib.call("bpmi_bp_polar_ex", _vp(d.data.ptr), d.data.shape[1], _vp(d.pos64.ptr),
                      _vp(d.weight64.ptr), _vp(d.phase64.ptr), begin, finish, _gp(geo), plan.nr,
                      plan.nu, plan.side, range_start, data_dr, k, _lib.mode_code(mode),
                      _lib.interp_code(self.interp), flags, _vp(accumulator.ptr),
                      stream_handle(stream), oref, -1)


CudaBaseline = HipBaseline


def _device_count():
    from .device import device_count
    return device_count()


_KAISER_READY = False


def kaiser_table():
    """Kaiser(beta=6)-windowed sinc at x = i/2048, i = 0..8193, FP64."""
    x = np.arange(_lib.KAISER_TABLE_LEN, dtype=np.float64) / 2048.0
    a = np.clip(x / 4.0, 0.0, 1.0)
    win = np.i0(6.0 * np.sqrt(1.0 - a * a)) / np.i0(6.0)
    val = np.sinc(x) * win
    val[x >= 4.0] = 0.0
    return val


def ensure_kaiser_table():
    global _KAISER_READY
    if not _KAISER_READY:
        table = kaiser_table()
        _lib.call("bpmi_set_kaiser_table", table.ctypes.data_as(_vp), len(table))
        _KAISER_READY = True


def hip_environment():
    """Device and toolchain report (counterpart of cuda_environment)."""
    info = device_info(0)
    free = _lib.c_size_t()
    total = _lib.c_size_t()
    _lib.call("bpmi_mem_info", _lib.ctypes.byref(free), _lib.ctypes.byref(total))
    return {
        "gpu": info["name"], "arch": info["arch"], "compute_units": info["compute_units"],
        "clock_khz": info["clock_khz"], "gpu_memory_bytes": int(total.value),
        "gpu_free_bytes_at_query": int(free.value),
        "hip_runtime_version": info["hip_runtime_version"],
        "numpy_version": np.__version__, "python_version": platform.python_version(),
        "platform": platform.platform(), "kernel_backend": "libbpmi (hipcc, gfx950)",
        "library": str(_lib.LIB_PATH), "fast_math": False,
    }


cuda_environment = hip_environment
