"""Phase-LUT direct-BP program (counterpart of fast_bp/resident_precision.py).

``make_lut_program`` keeps the reference's signature and returns an
uncaptured ``ResidentProgram`` whose schedule runs the LUT variant of mixed
direct BP: ``exp(j k (d - r) + j phase) = LUT[i] * exp(j (k dr f + phase)) *
exp(-j k r)`` with ``LUT[i] = exp(j k (range_start + i dr))`` generated once
in FP64 and stored complex64 (fast_bp/gpu/precision_probe.cu:100-164).
"""
from __future__ import annotations

import time

import numpy as np

from . import _lib
from .device import DeviceArray, Event, current_stream, elapsed_ms, stream_handle
from .resident_backend import ResidentProgram, _Launcher

_vp = _lib.c_void_p
_LUT_KERNEL = "direct_chunk_mixed_phase_lut"


def _attach_lut_to_schedule(program, lut):
    """Replace every mixed direct chunk of ``program`` with the LUT variant.

    Pure Python (no device work), so it can be exercised with a mock program
    that has ``_graph``, ``_schedule`` and ``operations`` attributes.  Must run
    before graph capture.
    """
    if program._graph is not None:
        raise RuntimeError("append LUT arguments before graph capture")
    if len(program._schedule) != len(program.operations):
        raise RuntimeError("resident schedule and operation list differ")
    for entry, operation in zip(program._schedule, program.operations):
        name, kernel, grid, block, args = entry
        if name != "direct_chunk_mixed" or operation["name"] != name:
            raise RuntimeError("expected an exclusively mixed direct-BP schedule")
        if operation["kernel"] is not kernel or operation["args"] is not args:
            raise RuntimeError("resident schedule and operation arguments differ")
    launcher = _Launcher("bpmi_bp_polar_lut")
    for index, (entry, operation) in enumerate(zip(program._schedule, program.operations)):
        _, kernel, grid, block, args = entry
        # bpmi_bp_polar_ex args: data, samples, pos, weight, phase, begin, end, geo, nr, nu,
        # side, range_start, data_dr, k, mode, interp, flags, out
        lut_args = args[:14] + (_vp(lut.ptr), args[17])
        program._schedule[index] = (_LUT_KERNEL, launcher, grid, block, lut_args)
        operation.update(name=_LUT_KERNEL, kernel=launcher, args=lut_args)


def make_lut_program(backend, plan, device, range_start, data_dr, k, chunk_pulses=256):
    """Reusable FP64-geometry, FP32-phase-LUT direct BP program."""
    if int(chunk_pulses) != chunk_pulses or chunk_pulses <= 0:
        raise ValueError("chunk_pulses must be a positive integer")
    if not all(np.isfinite(float(value)) for value in (range_start, data_dr, k)):
        raise ValueError("range_start, data_dr and k must be finite")
    if data_dr <= 0:
        raise ValueError("data_dr must be positive")
    samples = int(device.data.shape[1])
    if samples <= 0:
        raise ValueError("range data must have at least one sample")
    current_stream().synchronize()
    prepare_start = time.perf_counter()
    lut = DeviceArray.empty(samples, np.complex64)
    start, stop = Event(), Event()
    start.record()
    _lib.call("bpmi_fill_range_phase_lut", samples, float(range_start), float(data_dr), float(k),
              _vp(lut.ptr), stream_handle())
    stop.record()
    stop.synchronize()
    prepare_ms = (time.perf_counter() - prepare_start) * 1000.0
    program = ResidentProgram(backend, plan, device, "mixed", range_start, data_dr, k,
                              algorithm="direct", chunk_pulses=int(chunk_pulses))
    _attach_lut_to_schedule(program, lut)
    program._buffers.append(lut)
    lut_bytes = int(lut.nbytes)
    program.metadata.update({
        "variant": "phase_lut",
        "kernel_counts": {_LUT_KERNEL: program.launch_count},
        "arithmetic": ("FP64 geometry and range coordinates; four-point FP32 Lagrange; "
                       "FP64-generated complex64 LUT; FP32 sincosf fractional phase; "
                       "FP32 accumulation"),
        "lut_bytes": lut_bytes, "lut_entries": samples, "lut_dtype": "complex64",
        "lut_dependency_key": {"samples": samples, "range_start": float(range_start),
                               "data_dr": float(data_dr), "k": float(k)},
        "prepare_ms": float(prepare_ms),
        "lut_generation_device_ms": elapsed_ms(start, stop),
        "workspace_bytes": int(program.metadata["workspace_bytes"]) + lut_bytes,
        "workspace_buffer_count": len(program._buffers),
        "total_owned_device_bytes": int(program.metadata["total_owned_device_bytes"]) + lut_bytes,
    })
    return program
