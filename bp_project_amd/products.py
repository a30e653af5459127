"""Image products of an FFBP run: the Cartesian ground plane, the sampled
direct-BP validation and the files the reference writes
(fast_bp/run.py:90-128), with the heavy parts on the GPU.

* ``ground_image``  -- polar root -> uniform Cartesian ground grid on z = 0,
  remodulated (``evaluate_image`` semantics, fast_bp/ffbp.py:189-198), in one
  ``bpmi_eval_grid`` launch instead of run.py's 128-row host loop; rows are
  along-track ``y``, columns cross-track ``side * x`` as run.py:118-121 lays
  them out;
* ``sampled_direct_validation`` -- run.py:90-106: 3000 seeded interior
  pixels plus the 1000 brightest pixels of a thumbnail, re-imaged by direct
  BP at the same points (``bpmi_bp_points``) and compared;
* ``save_products`` -- run.py:107-128's files: ``{name}_validation_samples.npz``,
  ``{name}_metrics.json``, ``{name}_complex_demodulated.npy``,
  ``{name}_axes.npz`` and ``{name}_ground.mat`` (scipy ``savemat``), same
  keys.  Capture metadata of a real acquisition is the caller's to pass.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from . import _lib
from .device import DeviceArray

_vp = _lib.c_void_p


def _uniform_axis(v, name):
    v = np.asarray(v, np.float64).ravel()
    if len(v) == 0:
        raise ValueError(f"{name} axis is empty")
    if len(v) == 1:
        return float(v[0]), 1.0, 1
    step = (v[-1] - v[0]) / (len(v) - 1)
    if not np.allclose(np.diff(v), step, rtol=1e-9, atol=1e-9 * max(1.0, abs(step))):
        raise ValueError(f"{name} axis must be uniformly spaced")
    return float(v[0]), float(step), len(v)


def ground_image(result, x, y, k, side=None, *, mode=None, interp="lagrange4"):
    """Cartesian ground image of a ``PolarImage`` on the uniform axes x
    (cross-track distance, m) and y (along-track, m): complex64 [len(y), len(x)],
    pixel (a, b) at (side * x[b], y[a], 0), remodulated by exp(+j k R)."""
    from .ffbp import _interp, _mode
    side = result.side if side is None else side
    x0, dx, nx = _uniform_axis(x, "x")
    y0, dy, ny = _uniform_axis(y, "y")
    img = np.ascontiguousarray(result.image, np.complex64)
    src = DeviceArray.from_host(img)
    out = DeviceArray.empty((ny, nx))
    geo = np.array([*np.asarray(result.origin, np.float64), result.r0, result.dr, result.u0,
                    result.du], np.float64)
    _lib.call("bpmi_eval_grid", _vp(src.ptr), geo.ctypes.data_as(_vp), img.shape[0], img.shape[1],
              img.shape[1], side * x0, side * dx, nx, y0, dy, ny, float(k),
              _lib.mode_code(_mode(mode)), _interp(interp), _vp(out.ptr), None)
    return out.get()


def sample_pixels(result, seed=20180814, n_random=3000, n_bright=1000, guard=8):
    """Pixel indices of run.py:91-96: seeded interior pixels (``guard`` cells
    from every edge) plus the brightest pixels of a strided thumbnail."""
    nr, nu = result.image.shape
    rng = np.random.default_rng(seed)
    ir = rng.integers(guard, nr - guard, n_random)
    iu = rng.integers(guard, nu - guard, n_random)
    rstep = max(1, nr // 1800)
    ustep = max(1, nu // 1500)
    thumb = np.abs(result.image[::rstep, ::ustep])
    n_bright = min(n_bright, thumb.size)
    top = np.argpartition(thumb.ravel(), -n_bright)[-n_bright:]
    ri, ui = np.unravel_index(top, thumb.shape)
    return np.r_[ir, ri * rstep], np.r_[iu, ui * ustep]


def sampled_direct_validation(result, data, pos, r_start, data_dr, k, weight, phase, *,
                              seed=20180814, mode=None, interp="lagrange4"):
    """FFBP vs direct BP on sampled pixels (run.py:90-106).  Returns
    (validation dict, samples dict with ``direct``, ``ffbp``, ``xyz``)."""
    import time
    from .ffbp import backproject_points
    ir, iu = sample_pixels(result, seed=seed)
    rr = result.r[ir]
    uu = result.u[iu]
    o = np.asarray(result.origin, np.float64)
    ground = np.sqrt(rr ** 2 - o[2] ** 2)
    xyz = np.column_stack([o[0] + result.side * ground * np.sqrt(1 - uu * uu),
                           o[1] + ground * uu, np.zeros(len(rr))])
    t0 = time.perf_counter()
    direct = backproject_points(data, pos, xyz, r_start, data_dr, k, weight, phase, mode=mode,
                                interp=interp)
    seconds = time.perf_counter() - t0
    fast = result.image[ir, iu] * np.exp(1j * k * rr)
    error = float(np.linalg.norm(fast - direct) / np.linalg.norm(direct))
    coherence = float(abs(np.vdot(fast, direct))
                      / (np.linalg.norm(fast) * np.linalg.norm(direct)))
    validation = dict(points=len(rr), relative_complex_error=error, complex_coherence=coherence,
                      direct_bp_sample_seconds=seconds,
                      comparison="identical data, geometry, aperture weights and pixels",
                      passed=bool(error < .15 and coherence > .99))
    return validation, dict(direct=direct, ffbp=fast, xyz=xyz)


def save_products(out_dir, name, result, k, *, counts=None, validation=None, samples=None,
                  ground=None, x=None, y=None, metadata=None):
    """Write run.py:107-128's files for one image.  ``ground`` (with its axes
    ``x`` and ``y``) adds ``{name}_ground.mat``; ``metadata`` (capture,
    geometry, options, carrier_hz, height_reference_m, ...) goes into the
    metrics JSON and the .mat file.  Returns the written paths."""
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    meta = dict(metadata or {})
    written = []
    if samples is not None:
        p = out / f"{name}_validation_samples.npz"
        np.savez(p, direct=samples["direct"], ffbp=samples["ffbp"], xyz=samples["xyz"])
        written.append(p)
    metrics = {key: meta[key] for key in ("capture", "geometry", "options") if key in meta}
    metrics["ffbp"] = counts or {}
    metrics["validation"] = validation or {}
    p = out / f"{name}_metrics.json"
    p.write_text(json.dumps(metrics, indent=2, ensure_ascii=False, default=_jsonable),
                 encoding="utf-8")
    written.append(p)
    p = out / f"{name}_complex_demodulated.npy"
    np.save(p, np.asarray(result.image))
    written.append(p)
    p = out / f"{name}_axes.npz"
    np.savez(p, r=result.r, u=result.u, origin=np.asarray(result.origin), k=k, side=result.side)
    written.append(p)
    if ground is not None:
        from scipy.io import savemat
        if x is None or y is None:
            raise ValueError("ground image needs its x and y axes")
        p = out / f"{name}_ground.mat"
        mat = dict(image_complex=np.asarray(ground, np.complex64),
                   cross_track_m=result.side * np.asarray(x, np.float64),
                   along_track_m=np.asarray(y, np.float64),
                   phase_centre=np.asarray(result.origin, np.float64))
        for key in ("height_reference_m", "carrier_hz", "phase_sign", "valid_pulse_count"):
            if key in meta:
                mat[key] = meta[key]
        savemat(p, mat, do_compression=True)
        written.append(p)
    return written


def _jsonable(v):
    if isinstance(v, np.generic):
        return v.item()
    if isinstance(v, np.ndarray):
        return v.tolist()
    return str(v)
