"""The function set of fast_bp/ffbp.py, executed on the GPU through libbpmi.

Same names, arguments and return types as the reference's Numba module
(host numpy arrays in, complex64 numpy arrays out), so ``from
bp_project_amd.ffbp import form_ffbp, backproject_points, ...`` replaces
``from ffbp import ...``.  Each call uploads its inputs, runs the HIP
kernels and downloads the result (a PCIe-inclusive convenience path); the
device-resident path is ``gpu_backend`` / ``batched_resident_backend``.

Every function takes an optional ``mode`` (default ``DEFAULT_MODE = "fp64"``,
the arithmetic of the reference's CPU module: FP64 with complex64 stores)
and ``interp`` keyword; these extend, never change, the reference signature.
``cubic_weights``, ``sample1`` and ``sample2`` are the scalar helpers of the
reference, kept as small host functions for callers that use them directly.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .batched_resident_backend import BatchedResidentProgram
from .device import DeviceArray, synchronize
from .gpu_backend import ALL_MODES, HipBaseline, build_ffbp_plan, ensure_kaiser_table, geo_array

DEFAULT_MODE = "fp64"
C = 299792458.0
_vp = _lib.c_void_p
_backend = None


def _get_backend():
    global _backend
    if _backend is None:
        _backend = HipBaseline()
    return _backend


def _mode(mode):
    mode = DEFAULT_MODE if mode is None else mode
    if mode not in ALL_MODES:
        raise ValueError(f"mode must be one of {ALL_MODES}")
    return mode


def _interp(interp):
    code = _lib.interp_code(interp)
    if interp == "kaiser8":
        ensure_kaiser_table()
    return code


# ---------------------------------------------------------------------------
# scalar helpers (fast_bp/ffbp.py:17-40)
# ---------------------------------------------------------------------------

def cubic_weights(t):
    """Four-point Lagrange weights at taps floor(x)-1 .. floor(x)+2."""
    return (-t * (t - 1) * (t - 2) / 6, (t + 1) * (t - 1) * (t - 2) / 2,
            -(t + 1) * t * (t - 2) / 2, (t + 1) * t * (t - 1) / 6)


def sample1(row, q):
    i = int(math.floor(q))
    t = q - i
    if i < 1 or i + 2 >= len(row):
        return 0j
    w = cubic_weights(t)
    return w[0] * row[i - 1] + w[1] * row[i] + w[2] * row[i + 1] + w[3] * row[i + 2]


def sample2(img, r, u):
    ir = int(math.floor(r))
    iu = int(math.floor(u))
    if ir < 1 or iu < 1 or ir + 2 >= img.shape[0] or iu + 2 >= img.shape[1]:
        return 0j
    wr = cubic_weights(r - ir)
    wu = cubic_weights(u - iu)
    z = 0j
    for a in range(4):
        for b in range(4):
            z += img[ir + a - 1, iu + b - 1] * wr[a] * wu[b]
    return z


# ---------------------------------------------------------------------------
# kernels
# ---------------------------------------------------------------------------

def _c64(data):
    data = np.ascontiguousarray(data)
    if data.dtype != np.complex64:
        data = data.astype(np.complex64)
    return data


def _pulse_inputs(data, pos, weight, phase):
    data = _c64(data)
    P = data.shape[0]
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(P, 3)
    weight = np.ascontiguousarray(weight, dtype=np.float64).reshape(P)
    phase = np.ascontiguousarray(phase, dtype=np.float64).reshape(P)
    return (DeviceArray.from_host(data), DeviceArray.from_host(pos),
            DeviceArray.from_host(weight), DeviceArray.from_host(phase))


def backproject_points(data, pos, xyz, r_start, dr, k, weight, phase, *, mode=None,
                       interp="lagrange4"):
    """Direct BP at arbitrary xyz, not demodulated (fast_bp/ffbp.py:43-55)."""
    mode = _mode(mode)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    if len(xyz) == 0:
        return np.zeros(0, np.complex64)
    d, p, w, ph = _pulse_inputs(data, pos, weight, phase)
    pts = DeviceArray.from_host(xyz)
    out = DeviceArray.empty(len(xyz))
    _lib.call("bpmi_bp_points", _vp(d.ptr), d.shape[1], _vp(p.ptr), _vp(w.ptr), _vp(ph.ptr), 0,
              d.shape[0], _vp(pts.ptr), len(xyz), float(r_start), float(dr), float(k),
              _lib.mode_code(mode), _interp(interp), 0, _vp(out.ptr), None)
    return out.get()


def bp_polar(data, pos, origin, r0, dr, nr, u0, du, nu, r_start, data_dr, k, weight, phase, side,
             *, mode=None, interp="lagrange4"):
    """Polar BP demodulated by exp(-j k r) (fast_bp/ffbp.py:58-76)."""
    mode = _mode(mode)
    d, p, w, ph = _pulse_inputs(data, pos, weight, phase)
    out = DeviceArray.empty((int(nr), int(nu)))
    geo = np.array([origin[0], origin[1], origin[2], r0, dr, u0, du], dtype=np.float64)
    _lib.call("bpmi_bp_polar", _vp(d.ptr), d.shape[1], _vp(p.ptr), _vp(w.ptr), _vp(ph.ptr), 0,
              d.shape[0], geo.ctypes.data_as(_vp), int(nr), int(nu), int(side), float(r_start),
              float(data_dr), float(k), _lib.mode_code(mode), _interp(interp), 0, _vp(out.ptr),
              None)
    return out.get()


def regrid(img, src_origin, sr0, sdr, su0, sdu, dst_origin, r0, dr, nr, u0, du, nu, k, side, *,
           mode=None, interp="lagrange4"):
    """Child polar image -> parent polar grid (fast_bp/ffbp.py:79-95)."""
    mode = _mode(mode)
    img = _c64(img)
    src = DeviceArray.from_host(img)
    out = DeviceArray.empty((int(nr), int(nu)))
    sg = np.array([src_origin[0], src_origin[1], src_origin[2], sr0, sdr, su0, sdu], np.float64)
    dg = np.array([dst_origin[0], dst_origin[1], dst_origin[2], r0, dr, u0, du], np.float64)
    _lib.call("bpmi_regrid", _vp(src.ptr), sg.ctypes.data_as(_vp), img.shape[0], img.shape[1],
              dg.ctypes.data_as(_vp), int(nr), int(nu), int(side), float(k),
              _lib.mode_code(mode), _interp(interp), 0, _vp(out.ptr), None)
    return out.get()


@dataclass
class PolarImage:
    """fast_bp/ffbp.py:98-111."""
    image: np.ndarray
    origin: np.ndarray
    r0: float
    dr: float
    u0: float
    du: float
    side: int

    @property
    def r(self):
        return self.r0 + self.dr * np.arange(self.image.shape[0])

    @property
    def u(self):
        return self.u0 + self.du * np.arange(self.image.shape[1])


def form_ffbp(data, pos, r_start, data_dr, fc, extent, side=1, leaf_size=128, range_spacing=.25,
              angular_oversample=2.5, weight=None, phase=None, verbose=True, *, mode=None,
              interp="lagrange4", merge_interp="lagrange4"):
    """FFBP of a fixed world sector (fast_bp/ffbp.py:114-180), level-batched on
    the GPU.  Returns ``(PolarImage, counts)`` like the reference."""
    start = time.perf_counter()
    mode = _mode(mode)
    k = 4 * np.pi * fc / C
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    if weight is None:
        weight = np.ones(len(pos))
    if phase is None:
        phase = np.zeros(len(pos))
    plan = build_ffbp_plan(pos, fc, extent, side=side, leaf_size=leaf_size,
                           range_spacing=range_spacing, angular_oversample=angular_oversample)
    backend = HipBaseline(interp=interp, merge_interp=merge_interp)
    device = backend.upload(_c64(data), pos, weight, phase, plan)
    program = BatchedResidentProgram(backend, plan, device, mode, r_start, data_dr, k)
    program.execute()
    image = program.output.get()
    program.close()
    counts = {key: plan.counts[key] for key in
              ("leaves", "merges", "leaf_pixel_pulse_products", "merged_pixels")}
    counts["elapsed_s"] = time.perf_counter() - start
    counts["shape"] = list(image.shape)
    if verbose:
        print(f"FFBP: {counts['leaves']} leaves, {counts['merges']} merges, "
              f"{counts['elapsed_s']:.2f} s", flush=True)
    return PolarImage(image, plan.origin.copy(), plan.r0, plan.dr, plan.u0, plan.du, side), counts


def polar_points(origin, r, u, side=1):
    """fast_bp/ffbp.py:183-186."""
    rr, uu = np.meshgrid(r, u, indexing="ij")
    gr = np.sqrt(rr ** 2 - origin[2] ** 2)
    return np.column_stack([(origin[0] + side * gr * np.sqrt(1 - uu ** 2)).ravel(),
                            (origin[1] + gr * uu).ravel(), np.zeros(rr.size)])


def evaluate_image(img, origin, r0, dr, u0, du, xyz, k, *, mode=None, interp="lagrange4"):
    """Polar image -> points, remodulated by exp(+j k r) (fast_bp/ffbp.py:189-198)."""
    mode = _mode(mode)
    img = _c64(img)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    if len(xyz) == 0:
        return np.zeros(0, np.complex64)
    src = DeviceArray.from_host(img)
    pts = DeviceArray.from_host(xyz)
    out = DeviceArray.empty(len(xyz))
    geo = np.array([origin[0], origin[1], origin[2], r0, dr, u0, du], np.float64)
    _lib.call("bpmi_eval_points", _vp(src.ptr), geo.ctypes.data_as(_vp), img.shape[0],
              img.shape[1], img.shape[1], _vp(pts.ptr), len(xyz), float(k),
              _lib.mode_code(mode), _interp(interp), _vp(out.ptr), None)
    return out.get()


def target_contributions(data, pos, xyz, r_start, dr, k, *, mode=None, interp="lagrange4"):
    """Per-(point, pulse) unweighted contributions [M, P] (fast_bp/ffbp.py:201-211)."""
    mode = _mode(mode)
    data = _c64(data)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    P = data.shape[0]
    d = DeviceArray.from_host(data)
    p = DeviceArray.from_host(np.ascontiguousarray(pos, dtype=np.float64).reshape(P, 3))
    pts = DeviceArray.from_host(xyz)
    out = DeviceArray.empty((len(xyz), P))
    _lib.call("bpmi_target_contributions", _vp(d.ptr), d.shape[1], _vp(p.ptr), P, _vp(pts.ptr),
              len(xyz), float(r_start), float(dr), float(k), _lib.mode_code(mode),
              _interp(interp), _vp(out.ptr), None)
    return out.get()
