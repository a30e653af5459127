"""The drop-in boundary's errors, verbatim, raised before any device use.

The reference's messages (fast_bp/gpu_backend.py:199,247,274,279,281,443,
482-484; fast_bp/resident_backend.py:39-47,121,240,257,297-299;
fast_bp/resident_precision.py:85-92): the mode message lists the accepted
tuple, the reference's three modes followed by this build's dfloat and
split, in the same words from every entry point.
"""
import re
from types import SimpleNamespace

import numpy as np
import pytest

from bp_project_amd import _lib, ffbp
from bp_project_amd.batched_resident_backend import BatchedResidentProgram
from bp_project_amd.gpu_backend import ALL_MODES, MODES, HipBaseline, build_ffbp_plan, check_mode
from bp_project_amd.resident_backend import ResidentProgram

MODE_MSG = "mode must be one of ('fp64', 'fp32', 'mixed', 'dfloat', 'split')"


class Dummy:
    block_size = 128


def test_mode_tuple_extends_the_reference_tuple():
    assert MODES == ("fp64", "fp32", "mixed")
    assert ALL_MODES[:3] == MODES and ALL_MODES[3:] == ("dfloat", "split")


@pytest.mark.parametrize("call", [
    lambda: check_mode("fp16"),
    lambda: _lib.mode_code("fp16"),
    lambda: ffbp._mode("fp16"),
    lambda: ResidentProgram(Dummy(), None, None, "fp16", 0, 1, 1),
    lambda: BatchedResidentProgram(Dummy(), None, None, "fp16", 0, 1, 1),
    lambda: object.__new__(HipBaseline).run_direct(None, None, "fp16", 0, 1, 1),
    lambda: object.__new__(HipBaseline).run_ffbp(None, None, "fp16", 0, 1, 1),
])
def test_mode_error_is_the_same_everywhere(call):
    with pytest.raises(ValueError, match=f"^{re.escape(MODE_MSG)}$"):
        call()


def _exact(exc, msg, call):
    with pytest.raises(exc, match=f"^{re.escape(msg)}$"):
        call()


def test_argument_errors_verbatim():
    bare = object.__new__(HipBaseline)
    _exact(ValueError, "chunk_pulses must be positive",
           lambda: bare.run_direct(None, None, "fp64", 0, 1, 1, chunk_pulses=0))
    _exact(ValueError, "chunk_pulses must be positive",
           lambda: ResidentProgram(Dummy(), None, None, "fp64", 0, 1, 1, chunk_pulses=0))
    _exact(ValueError, "block_size must be positive", lambda: HipBaseline(block_size=0))
    msg = "range-compressed data must be C-contiguous complex64 [pulse, range]"
    _exact(TypeError, msg, lambda: bare.upload(np.zeros((4, 8), np.complex128), np.zeros((4, 3)),
                                               np.ones(4), np.zeros(4), None))
    _exact(TypeError, msg, lambda: bare.upload(np.zeros(8, np.complex64), np.zeros((8, 3)),
                                               np.ones(8), np.zeros(8), None))
    data = np.zeros((4, 8), np.complex64)
    _exact(ValueError, "position shape does not match the pulse count",
           lambda: bare.upload(data, np.zeros((5, 3)), np.ones(4), np.zeros(4), None))
    _exact(ValueError, "weight and phase must have one value per pulse",
           lambda: bare.upload(data, np.zeros((4, 3)), np.ones(3), np.zeros(4), None))
    _exact(ValueError, "algorithm must be direct or ffbp",
           lambda: ResidentProgram(Dummy(), None, None, "fp64", 0, 1, 1, algorithm="x"))
    _exact(ValueError, "variant must be original or fused",
           lambda: ResidentProgram(Dummy(), None, None, "fp64", 0, 1, 1, variant="x"))
    _exact(ValueError, "fused variant is only defined for FFBP",
           lambda: ResidentProgram(Dummy(), None, None, "fp64", 0, 1, 1, algorithm="direct",
                                   variant="fused"))


def test_extent_and_empty_grid_errors():
    from bp_project_amd.scenes import synthetic_kat_scene
    c = synthetic_kat_scene()
    # an extent narrower than one range cell between two grid rows
    with pytest.raises(ValueError,
                       match="^Requested extent does not intersect the planned root grid$"):
        build_ffbp_plan(c["pos"], c["fc"], (730.05, 730.06, -.2, -.19), leaf_size=128,
                        range_spacing=.5, angular_oversample=3.)
    prog = object.__new__(ResidentProgram)
    prog.backend = Dummy()
    _exact(ValueError, "empty image grids are unsupported", lambda: prog._enqueue("x", 0, ()))


def test_program_state_errors_verbatim():
    prog = object.__new__(ResidentProgram)
    prog._closed, prog._graph = True, None
    for method in (prog.execute, prog.capture_graph, prog.replay):
        _exact(RuntimeError, "program is closed", method)
    prog._closed = False
    _exact(RuntimeError, "capture_graph() must succeed before replay()", prog.replay)


def test_lut_program_argument_errors_verbatim():
    from bp_project_amd.resident_precision import make_lut_program
    plan = SimpleNamespace(nr=4, nu=4)
    dev = SimpleNamespace(data=SimpleNamespace(shape=(4, 16)))
    _exact(ValueError, "chunk_pulses must be a positive integer",
           lambda: make_lut_program(Dummy(), plan, dev, 0.0, 1.0, 1.0, chunk_pulses=0))
    _exact(ValueError, "range_start, data_dr and k must be finite",
           lambda: make_lut_program(Dummy(), plan, dev, np.nan, 1.0, 1.0))
    _exact(ValueError, "data_dr must be positive",
           lambda: make_lut_program(Dummy(), plan, dev, 0.0, -1.0, 1.0))
    dev0 = SimpleNamespace(data=SimpleNamespace(shape=(4, 0)))
    _exact(ValueError, "range data must have at least one sample",
           lambda: make_lut_program(Dummy(), plan, dev0, 0.0, 1.0, 1.0))
