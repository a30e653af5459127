"""ctypes binding of libbpmi.so (the C ABI declared in include/bpmi.h).

The library is loaded from the package directory only.  There is no CPU
fallback: if the shared object is missing or no HIP device is present, every
entry point raises.  Status codes map to the reference's exception types:
BPMI_ERR_INVALID -> ValueError, anything else -> RuntimeError.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import c_double, c_float, c_int, c_longlong, c_size_t, c_void_p, c_char_p, POINTER
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libbpmi.so"

OK, ERR_INVALID, ERR_HIP, ERR_UNSUPPORTED = 0, 1, 2, 3
MODE_FP64, MODE_FP32, MODE_MIXED, MODE_DFLOAT, MODE_SPLIT = 0, 1, 2, 3, 4
MODE_CODES = {"fp64": MODE_FP64, "fp32": MODE_FP32, "mixed": MODE_MIXED,
              "dfloat": MODE_DFLOAT, "split": MODE_SPLIT}
INTERP_CODES = {"lagrange4": 0, "linear": 1, "kaiser8": 2}
ACCUMULATE, OUT_C128 = 1, 2
KAISER_TABLE_LEN = 4 * 2048 + 2
# Launch shapes accepted by bpmi_ffbp_leaves_multi_ex (split-kernel pulse
# block depth) and bpmi_ffbp_merges_tiled_ex ((waves per SIMD hint, window
# cap)); the first of each is the shape of the entry points without _ex.
PULSE_BLOCKS = (32, 16, 64)
MERGE_SHAPES = ((4, 2176), (5, 1664), (5, 1920), (7, 1088))
DEFAULT_PULSE_BLOCK = PULSE_BLOCKS[0]
DEFAULT_MERGE_SHAPE = MERGE_SHAPES[0]

_vp, _i, _d, _ll, _sz = c_void_p, c_int, c_double, c_longlong, c_size_t

# name -> argument types (every function returns int status)
SIGNATURES = {
    "bpmi_version": [],
    "bpmi_last_error": [c_char_p, _sz],
    "bpmi_device_count": [POINTER(c_int)],
    "bpmi_set_device": [_i],
    "bpmi_get_device": [POINTER(c_int)],
    "bpmi_device_info": [_i, c_char_p, _sz, c_char_p, _sz, POINTER(c_int), POINTER(c_int),
                         POINTER(c_size_t)],
    "bpmi_runtime_version": [POINTER(c_int)],
    "bpmi_mem_info": [POINTER(c_size_t), POINTER(c_size_t)],
    "bpmi_wall_clock_khz": [POINTER(c_int)],
    "bpmi_malloc": [POINTER(c_void_p), _sz],
    "bpmi_free": [_vp],
    "bpmi_host_alloc": [POINTER(c_void_p), _sz],
    "bpmi_host_free": [_vp],
    "bpmi_memcpy_h2d": [_vp, _vp, _sz, _vp],
    "bpmi_memcpy_d2h": [_vp, _vp, _sz, _vp],
    "bpmi_memcpy_d2d": [_vp, _vp, _sz, _vp],
    "bpmi_zero": [_vp, _sz, _vp],
    "bpmi_stream_create": [POINTER(c_void_p)],
    "bpmi_stream_destroy": [_vp],
    "bpmi_stream_synchronize": [_vp],
    "bpmi_device_synchronize": [],
    "bpmi_event_create": [POINTER(c_void_p)],
    "bpmi_event_destroy": [_vp],
    "bpmi_event_record": [_vp, _vp],
    "bpmi_event_synchronize": [_vp],
    "bpmi_event_elapsed": [POINTER(c_float), _vp, _vp],
    "bpmi_graph_begin": [_vp],
    "bpmi_graph_end": [_vp, POINTER(c_void_p)],
    "bpmi_graph_launch": [_vp, _vp],
    "bpmi_graph_destroy": [_vp],
    "bpmi_bp_polar": [_vp, _i, _vp, _vp, _vp, _i, _i, _vp, _i, _i, _i, _d, _d, _d, _i, _i, _i,
                      _vp, _vp],
    "bpmi_bp_polar_ex": [_vp, _i, _vp, _vp, _vp, _i, _i, _vp, _i, _i, _i, _d, _d, _d, _i, _i, _i,
                         _vp, _vp, _i, _i],
    "bpmi_bp_grid": [_vp, _i, _vp, _vp, _vp, _i, _i, _d, _d, _i, _d, _d, _i, _d, _d, _d, _d, _i,
                     _i, _i, _vp, _vp],
    "bpmi_bp_grid_ws": [_vp, _i, _vp, _vp, _vp, _i, _i, _d, _d, _i, _d, _d, _i, _d, _d, _d, _d,
                        _i, _i, _i, _vp, _vp, _sz, _vp],
    "bpmi_bp_grid_workspace_bytes": [_i, _i, _i, POINTER(c_size_t)],
    "bpmi_bp_grid_noalloc": [_vp, _i, _vp, _vp, _vp, _i, _i, _d, _d, _i, _d, _d, _i, _d, _d, _d,
                             _d, _i, _i, _i, _vp, _vp],
    "bpmi_bp_points": [_vp, _i, _vp, _vp, _vp, _i, _i, _vp, _ll, _d, _d, _d, _i, _i, _i, _vp,
                       _vp],
    "bpmi_target_contributions": [_vp, _i, _vp, _i, _vp, _ll, _d, _d, _d, _i, _i, _vp, _vp],
    "bpmi_regrid": [_vp, _vp, _i, _i, _vp, _i, _i, _i, _d, _i, _i, _i, _vp, _vp],
    "bpmi_merge2": [_vp, _vp, _i, _i, _vp, _vp, _i, _i, _vp, _i, _i, _i, _d, _i, _i, _vp, _vp],
    "bpmi_add_inplace": [_vp, _vp, _ll, _vp],
    "bpmi_cast_c128_c64": [_vp, _vp, _ll, _vp],
    "bpmi_crop": [_vp, _i, _i, _i, _i, _i, _vp, _vp],
    "bpmi_eval_points": [_vp, _vp, _i, _i, _i, _vp, _ll, _d, _i, _i, _vp, _vp],
    "bpmi_eval_grid": [_vp, _vp, _i, _i, _i, _d, _d, _i, _d, _d, _i, _d, _i, _i, _vp, _vp],
    "bpmi_ffbp_leaves": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i, _d, _d, _d, _i, _i,
                         _vp],
    "bpmi_ffbp_merges": [_vp, _vp, _vp, _vp, _vp, _i, _d, _i, _i, _vp],
    "bpmi_split_tiles": [_i, _i, POINTER(c_int)],
    "bpmi_split_tiles_ex": [_i, _i, _i, POINTER(c_int)],
    "bpmi_ffbp_merges_tiled": [_vp, _vp, _vp, _vp, _vp, _i, _d, _i, _vp],
    "bpmi_ffbp_merges_tiled_ex": [_vp, _vp, _vp, _vp, _vp, _i, _d, _i, _i, _i, _vp],
    "bpmi_ffbp_merge_windows": [_vp, _vp, _vp, _i, _i, _vp, _vp],
    "bpmi_ffbp_merges_tiled_win": [_vp, _vp, _vp, _vp, _vp, _i, _d, _i, _i, _i, _vp, _vp],
    "bpmi_merge_tiles": [_i, _i, POINTER(c_int)],
    "bpmi_ffbp_required_tiles": [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "bpmi_eval_required_tiles": [_vp, _i, _i, _i, _i, _i, _i, _d, _d, _i, _d, _d, _i, _i, _vp,
                                 _vp],
    "bpmi_ffbp_leaves_multi": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _vp],
    "bpmi_ffbp_leaves_multi_ex": [_vp, _vp, _vp, _vp, _vp, _vp, _i, _i, _i, _i, _i, _vp],
    "bpmi_eval_grid_multi": [_vp, _vp, _i, _i, _i, _vp],
    "bpmi_eval_grid_multi_tiled": [_vp, _vp, _i, _i, _vp],
    "bpmi_eval_tiles": [_i, _i, POINTER(c_int)],
    "bpmi_fill_range_phase_lut": [_i, _d, _d, _d, _vp, _vp],
    "bpmi_bp_polar_lut": [_vp, _i, _vp, _vp, _vp, _i, _i, _vp, _i, _i, _i, _d, _d, _d, _vp, _vp,
                          _vp],
    "bpmi_set_kaiser_table": [_vp, _i],
    "bpmi_adjoint_workspace_bytes": [_ll, _i, POINTER(c_size_t)],
    "bpmi_grid_adjoint_workspace_bytes": [_i, _i, _i, POINTER(c_size_t)],
    "bpmi_bp_grid_adjoint": [_vp, _i, _vp, _vp, _vp, _i, _i, _d, _d, _i, _d, _d, _i, _d, _d, _d,
                             _d, _i, _vp, _vp, _sz, _vp, _vp, _vp],
    "bpmi_bp_points_adjoint": [_vp, _i, _vp, _vp, _vp, _i, _i, _vp, _ll, _d, _d, _d, _i, _vp, _vp,
                               _sz, _vp, _vp, _vp],
    "bpmi_probe_sincos_turns": [_vp, _vp, _vp, _ll, _vp],
    "bpmi_probe_fma_peak": [_vp, _i, _i, _vp],
    "bpmi_range_compress_workspace_bytes": [_i, _i, _d, _d, _i, _i, POINTER(c_size_t)],
    "bpmi_range_compress": [_vp, _i, _i, _ll, _d, _d, _d, _i, _i, _i, _vp, _vp, _sz, _vp],
    "bpmi_decode_iq16": [_vp, _ll, _ll, _ll, _i, _vp, _ll, _vp, _vp],
    "bpmi_persistent_workspace_bytes": [_i, _i, POINTER(c_size_t)],
    "bpmi_persistent_grid": [_i, _i, POINTER(c_int)],
    "bpmi_ffbp_persistent": [_vp, _i, _i, _i, _vp],
    "bpmi_persistent_aborted": [_vp, _i, _i, POINTER(c_int)],
    "bpmi_persistent_stamps": [_vp, _i, _i, _vp],
}


class PersistentSchedule(ctypes.Structure):
    """bpmi_persistent_schedule (include/bpmi.h), field for field."""
    _fields_ = [
        ("node_int", _vp), ("node_geo", _vp), ("offsets", _vp), ("arena", _vp),
        ("sources", _vp), ("data", _vp), ("pos", _vp), ("weight", _vp), ("phase", _vp),
        ("samples", _i), ("tile_cols_log2", _i),
        ("range_start", _d), ("data_dr", _d), ("k", _d),
        ("jobs", _vp), ("njobs", _i),
        ("items", _vp), ("nitems", _i),
        ("node_tiles", _vp), ("nodes", _i),
        ("job_root", _vp), ("job_blocks", _vp),
        ("workspace", _vp), ("max_wait_s", _d),
    ]

_lib = None


class BpmiError(RuntimeError):
    pass


def _share_hip_runtime_with_torch():
    """Import torch (when installed) before libbpmi.so is loaded.

    The torch wheel carries its own HIP runtime; loaded first, it is the one
    runtime of the process and libbpmi.so binds to it, so torch tensors,
    streams and libbpmi calls share one device context.  Loaded the other
    way round, torch finds no usable device.  ``BPMI_NO_TORCH=1`` skips the
    import for programs that never use torch.
    """
    if os.environ.get("BPMI_NO_TORCH") == "1":
        return
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load(path=None):
    """Load libbpmi.so (no GPU is touched by loading)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"libbpmi.so not found at {p}: build it with `python -m bp_project_amd.build` "
            "(there is no CPU fallback)")
    _share_hip_runtime_with_torch()
    lib = ctypes.CDLL(str(p))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = c_int
    if path is None:
        _lib = lib
    return lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().bpmi_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status, what=""):
    if status == OK:
        return
    msg = last_error()
    if status == ERR_INVALID:
        raise ValueError(msg)
    raise BpmiError(f"{what}: {msg}" if what else msg)


def call(name, *args):
    """Call a bpmi_* function and raise on a nonzero status."""
    check(getattr(load(), name)(*args), name)


def mode_code(mode):
    try:
        return MODE_CODES[mode]
    except KeyError:
        raise ValueError(f"mode must be one of {tuple(MODE_CODES)}") from None


def interp_code(interp):
    try:
        return INTERP_CODES[interp]
    except KeyError:
        raise ValueError(f"interp must be one of {tuple(INTERP_CODES)}") from None
