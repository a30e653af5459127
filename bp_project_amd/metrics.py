"""Image-quality metrics with the reference's conventions (host numpy code).

* ``fidelity(reference, estimate)`` -- relative complex error (NRMSE),
  SER, coherence, magnitude error, finite fraction and bias-removed phase
  statistics above -30 dB, peak shift (fast_bp/gpu_baseline.py:49-94).
* ``focus_metrics(z)`` -- entropy of normalised intensity, contrasts
  (sar_quality/metrics.py:6-22).
* ``impulse_metrics(x, z)`` -- 1-D half-power IRW, first-minima mainlobe,
  PSLR and ISLR (sar_quality/metrics.py:42-74).
* ``profile_metrics`` / ``analyze_target`` -- the qualified-minima 1-D/2-D
  point-response measurements (vancouver_bp/quality.py:73-239).
* ``peak_phase_error`` -- phase difference at each target's peak pixel.
* ``timing_summary`` -- median/IQR of timing samples.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.ndimage import gaussian_filter1d
from scipy.signal import find_peaks


def _trapz(y, x, axis=-1):
    return np.trapezoid(y, x, axis=axis)


def _db(value):
    return float(10.0 * np.log10(max(float(value), np.finfo(float).tiny)))


def fidelity(reference, estimate):
    reference = np.asarray(reference, dtype=np.complex128)
    estimate = np.asarray(estimate, dtype=np.complex128)
    finite = np.isfinite(estimate)
    ref_norm = np.linalg.norm(reference.ravel())
    est_norm = np.linalg.norm(estimate.ravel())
    error = np.linalg.norm((estimate - reference).ravel()) / ref_norm
    mask = np.abs(reference) >= np.abs(reference).max() * 10.0 ** (-30.0 / 20.0)
    mask &= finite
    phase = np.angle(estimate[mask] * np.conj(reference[mask]))
    if phase.size:
        bias = np.angle(np.mean(np.exp(1j * phase)))
        residual = np.angle(np.exp(1j * (phase - bias)))
        phase_std = np.rad2deg(residual.std())
        phase_median_abs = np.rad2deg(np.median(np.abs(residual)))
        phase_p95_abs = np.rad2deg(np.percentile(np.abs(residual), 95))
    else:
        bias = phase_std = phase_median_abs = phase_p95_abs = math.nan
    ref_peak = np.unravel_index(np.argmax(np.abs(reference)), reference.shape)
    est_peak = np.unravel_index(np.argmax(np.abs(estimate)), estimate.shape)
    peak_db = 20.0 * np.log10(max(float(np.abs(estimate).max()), 1e-300)
                              / max(float(np.abs(reference).max()), 1e-300))
    magnitude_error = (np.linalg.norm((np.abs(estimate) - np.abs(reference)).ravel())
                       / np.linalg.norm(np.abs(reference).ravel()))
    return {
        "relative_complex_error": float(error),
        "ser_db": float(-20.0 * np.log10(max(error, 1e-300))),
        "complex_coherence": float(abs(np.vdot(reference.ravel(), estimate.ravel()))
                                   / (ref_norm * est_norm)),
        "relative_magnitude_error": float(magnitude_error),
        "finite_fraction": float(finite.mean()),
        "phase_bias_deg": float(np.rad2deg(bias)),
        "phase_std_deg_above_minus30db": float(phase_std),
        "phase_median_abs_deg_above_minus30db": float(phase_median_abs),
        "phase_p95_abs_deg_above_minus30db": float(phase_p95_abs),
        "phase_mask_samples": int(mask.sum()),
        "peak_amplitude_difference_db": float(peak_db),
        "reference_peak_pixel": [int(x) for x in ref_peak],
        "estimate_peak_pixel": [int(x) for x in est_peak],
        "peak_shift_pixels": [int(est_peak[i] - ref_peak[i]) for i in range(len(ref_peak))],
    }


def timing_summary(values):
    values = np.asarray(values, dtype=np.float64)
    return {"samples_ms": [float(x) for x in values], "median_ms": float(np.median(values)),
            "q1_ms": float(np.percentile(values, 25)), "q3_ms": float(np.percentile(values, 75)),
            "min_ms": float(values.min()), "max_ms": float(values.max())}


def focus_metrics(z):
    z = np.asarray(z)
    valid = np.isfinite(z)
    out = dict(samples=int(z.size), finite_fraction=float(valid.mean()),
               nonzero_fraction=float(np.mean(abs(z) > 0)))
    if not valid.all() or not np.any(abs(z) > 0):
        return dict(out, valid=False, reason="Non-finite or zero-energy image")
    a = abs(z).astype(np.float64).ravel()
    a /= a.max()
    intensity = a * a
    p = intensity / intensity.sum()
    nz = p > 0
    entropy = float(-np.sum(p[nz] * np.log(p[nz])))
    return dict(out, valid=True, entropy_nats=entropy, normalized_entropy=entropy / np.log(z.size),
                intensity_contrast=float(intensity.std() / intensity.mean()),
                amplitude_contrast=float(a.std() / a.mean()),
                brightest_1pct_energy_fraction=float(np.sort(p)[-max(1, z.size // 100):].sum()))


def image_entropy(z):
    """Entropy (nats) of p = |z|^2 / sum|z|^2."""
    p = np.abs(np.asarray(z, np.complex128).ravel()) ** 2
    p = p / p.sum()
    nz = p > 0
    return float(-np.sum(p[nz] * np.log(p[nz])))


def impulse_metrics(x, z):
    """1-D half-power IRW; first-minima mainlobe; finite-cut PSLR/ISLR."""
    x = np.asarray(x, float)
    p = abs(np.asarray(z)) ** 2
    if not np.isfinite(p).all() or p.max() <= 0:
        return dict(valid=False, reason="Invalid power")
    peak = int(np.argmax(p))
    p = p / p[peak]
    left = np.flatnonzero(p[:peak] <= .5)
    right = np.flatnonzero(p[peak + 1:] <= .5) + peak + 1
    if not len(left) or not len(right):
        return dict(valid=False, reason="Half-power crossing absent")
    il, ir = int(left[-1]), int(right[0])
    xl = np.interp(.5, p[il:il + 2], x[il:il + 2])
    xr = np.interp(.5, p[ir - 1:ir + 1][::-1], x[ir - 1:ir + 1][::-1])
    mins = find_peaks(-p)[0]
    ml, mr = mins[mins < il], mins[mins > ir]
    if not len(ml) or not len(mr):
        return dict(valid=True, irw_m=float(xr - xl), peak_offset_m=float(x[peak]),
                    sidelobes_valid=False, reason="First minima absent")
    ml, mr = int(ml[-1]), int(mr[0])
    em = _trapz(p[ml:mr + 1], x[ml:mr + 1])
    es = _trapz(p[:ml + 1], x[:ml + 1]) + _trapz(p[mr:], x[mr:])
    return dict(valid=True, sidelobes_valid=True, irw_m=float(xr - xl),
                peak_offset_m=float(x[peak]),
                pslr_db=float(10 * np.log10(max(p[:ml].max(), p[mr + 1:].max()))),
                islr_db=float(10 * np.log10(es / em)),
                mainlobe_bounds_m=[float(x[ml]), float(x[mr])])


def _crossing(x0, x1, p0, p1, level=0.5):
    if p1 == p0:
        return float((x0 + x1) / 2)
    return float(x0 + (level - p0) * (x1 - x0) / (p1 - p0))


def _qualified_minima(x, p, il, ir, irw):
    dx = float(np.median(np.diff(x)))
    sigma = min(4.0, max(0.5, 0.04 * irw / dx))
    smooth = gaussian_filter1d(p, sigma, mode="nearest")
    minima = find_peaks(-smooth)[0]
    left = minima[(minima < il) & (minima > 0)][::-1]
    right = minima[(minima > ir) & (minima < len(p) - 1)]

    def first_deep(indices):
        for limit in (0.1, 0.25):
            deep = [int(i) for i in indices if p[i] <= limit]
            if deep:
                return deep[0]
        return None

    return first_deep(left), first_deep(right), sigma


def profile_metrics(x, z, *, expected_irw_m=None, integration_half_width_m=None, peak_index=None):
    """1-D point response with qualified first minima (vancouver_bp/quality.py:73-176)."""
    x = np.asarray(x, dtype=np.float64)
    z = np.asarray(z)
    if x.ndim != 1 or z.ndim != 1 or len(x) != len(z) or len(x) < 5:
        raise ValueError("x and z must be matching one-dimensional arrays with at least five samples")
    if not np.all(np.isfinite(x)) or np.any(np.diff(x) <= 0):
        raise ValueError("x must be finite and strictly increasing")
    power = np.abs(z.astype(np.complex128)) ** 2
    result = dict(valid=False, sidelobes_valid=False, warnings=[])
    if not np.all(np.isfinite(power)) or not np.any(power > 0):
        return dict(result, reason="nonfinite or zero power")
    peak_all = int(np.argmax(power)) if peak_index is None else int(peak_index)
    if not 0 <= peak_all < len(power) or power[peak_all] <= 0:
        raise ValueError("peak_index must identify a positive-power sample inside the supplied cut")
    centre = float(x[peak_all])
    truncated = False
    if integration_half_width_m is not None:
        hw = float(integration_half_width_m)
        if hw <= 0:
            raise ValueError("integration_half_width_m must be positive")
        truncated = centre - hw < x[0] or centre + hw > x[-1]
        mask = (x >= centre - hw) & (x <= centre + hw)
        idx = np.flatnonzero(mask)
        x, power = x[mask], power[mask]
    else:
        idx = np.arange(len(x))
    if len(x) < 5:
        return dict(result, reason="integration window contains fewer than five samples")
    peak = int(peak_all - idx[0])
    p = power / float(power[peak])
    dx = np.diff(x)
    result.update(peak_coordinate_m=float(x[peak]), peak_sample_index=int(idx[peak]),
                  step_m=float(np.median(dx)), integration_bounds_m=[float(x[0]), float(x[-1])],
                  integration_window_truncated=bool(truncated))
    below_left = np.flatnonzero(p[:peak] <= 0.5)
    below_right = np.flatnonzero(p[peak + 1:] <= 0.5) + peak + 1
    if not len(below_left) or not len(below_right):
        return dict(result, reason="half-power crossings absent on one or both sides")
    il, ir = int(below_left[-1]), int(below_right[0])
    xl = _crossing(x[il], x[il + 1], p[il], p[il + 1])
    xr = _crossing(x[ir - 1], x[ir], p[ir - 1], p[ir])
    irw = float(xr - xl)
    result.update(valid=True, irw_m=irw, half_power_bounds_m=[xl, xr])
    if expected_irw_m is not None:
        result["irw_to_expected_ratio"] = irw / float(expected_irw_m)
    ml, mr, _ = _qualified_minima(x, p, il, ir, irw)
    if ml is None or mr is None:
        return dict(result, reason="qualified first minima absent; sidelobe ratios not identified")
    if ml == 0 or mr == len(p) - 1:
        return dict(result, reason="mainlobe touches integration boundary")
    main = float(_trapz(p[ml:mr + 1], x[ml:mr + 1]))
    side = float(_trapz(p[:ml + 1], x[:ml + 1]) + _trapz(p[mr:], x[mr:]))
    side_values = np.r_[p[:ml], p[mr + 1:]]
    if main <= 0 or not len(side_values):
        return dict(result, reason="main/side integration regions invalid")
    depths = [_db(p[ml]), _db(p[mr])]
    half_components = int(np.count_nonzero(np.diff(np.r_[False, p >= .5, False].astype(int)) == 1))
    deep = bool(max(depths) <= -15.0)
    reliable = deep and not truncated and half_components == 1
    result.update(sidelobes_valid=bool(reliable), pslr_db=_db(float(np.max(side_values))),
                  islr_db=_db(side / main), mainlobe_bounds_m=[float(x[ml]), float(x[mr])],
                  mainlobe_minima_db=depths,
                  mainlobe_sample_indices=[int(idx[ml]), int(idx[mr])],
                  half_power_components=half_components)
    if not deep:
        result["warnings"].append("mainlobe minima shallower than -15 dB")
    if truncated:
        result["warnings"].append("requested integration window exceeds supplied cut")
    if half_components != 1:
        result["warnings"].append("multiple half-power components")
    return result


def analyze_target(r, y, image, *, peak_hint=None, search_radius_m=None,
                   integration_half_widths_m=None):
    """Range and azimuth cuts through one target of image[r, y]
    (vancouver_bp/quality.py:179-239, 1-D part)."""
    r, y = np.asarray(r, float), np.asarray(y, float)
    image = np.asarray(image)
    if image.shape != (len(r), len(y)):
        raise ValueError("image shape must be (len(r), len(y))")
    power = np.abs(image.astype(np.complex128)) ** 2
    if not np.all(np.isfinite(power)) or not np.any(power > 0):
        raise ValueError("target image has nonfinite or zero power")
    selection = np.ones(image.shape, bool)
    if peak_hint is not None and search_radius_m is not None:
        dr, dy = search_radius_m if np.ndim(search_radius_m) else (search_radius_m,) * 2
        selection = ((np.abs(r[:, None] - peak_hint[0]) <= dr)
                     & (np.abs(y[None, :] - peak_hint[1]) <= dy))
        if not np.any(selection):
            raise ValueError("peak search window contains no pixels")
    ir, iy = np.unravel_index(np.argmax(np.where(selection, power, -np.inf)), image.shape)
    half = (None, None) if integration_half_widths_m is None else integration_half_widths_m
    rng = profile_metrics(r, image[:, iy], integration_half_width_m=half[0], peak_index=ir)
    az = profile_metrics(y, image[ir, :], integration_half_width_m=half[1], peak_index=iy)
    return dict(peak=dict(range_m=float(r[ir]), azimuth_m=float(y[iy]), index=[int(ir), int(iy)],
                          amplitude=float(np.sqrt(power[ir, iy]))),
                range=rng, azimuth=az)


def peak_phase_error(reference, estimate, peaks):
    """|arg(est * conj(ref))| at each (row, col) peak, radians."""
    out = []
    for (i, j) in peaks:
        out.append(float(abs(np.angle(estimate[i, j] * np.conj(reference[i, j])))))
    return out


def target_peaks(image, grid, targets, search_m=2.0):
    """Peak pixel of |image| within +-search_m of each target on a Cartesian
    grid image [ny, nx]."""
    peaks = []
    amp = np.abs(image)
    for t in targets:
        ix = int(round((t[0] - grid.x0) / grid.dx))
        iy = int(round((t[1] - grid.y0) / grid.dy))
        h = max(1, int(round(search_m / abs(grid.dx))))
        y0, y1 = max(0, iy - h), min(grid.ny, iy + h + 1)
        x0, x1 = max(0, ix - h), min(grid.nx, ix + h + 1)
        if y0 >= y1 or x0 >= x1:
            continue
        sub = amp[y0:y1, x0:x1]
        a, b = np.unravel_index(np.argmax(sub), sub.shape)
        peaks.append((y0 + int(a), x0 + int(b)))
    return peaks


def pslr_islr_deltas(reference, estimate, grid, peaks, half_width_m=8.0):
    """PSLR/ISLR (dB) of range (x) and azimuth (y) cuts through each peak in
    ``estimate`` minus the same in ``reference``; returns the worst absolute
    deltas and per-target values."""
    x, y = grid.x, grid.y
    rows = []
    for (i, j) in peaks:
        rec = {}
        for name, ref_cut, est_cut, axis, idx in (
                ("range", reference[i, :], estimate[i, :], x, j),
                ("azimuth", reference[:, j], estimate[:, j], y, i)):
            ax = axis if axis[1] > axis[0] else axis[::-1]
            rc = ref_cut if axis[1] > axis[0] else ref_cut[::-1]
            ec = est_cut if axis[1] > axis[0] else est_cut[::-1]
            pk = idx if axis[1] > axis[0] else len(axis) - 1 - idx
            a = profile_metrics(ax, rc, integration_half_width_m=half_width_m, peak_index=pk)
            b = profile_metrics(ax, ec, integration_half_width_m=half_width_m, peak_index=pk)
            if "pslr_db" in a and "pslr_db" in b:
                rec[name] = dict(pslr_delta_db=b["pslr_db"] - a["pslr_db"],
                                 islr_delta_db=b["islr_db"] - a["islr_db"],
                                 pslr_ref_db=a["pslr_db"], islr_ref_db=a["islr_db"])
        rows.append(rec)
    deltas = [abs(v[k]) for rec in rows for v in rec.values()
              for k in ("pslr_delta_db", "islr_delta_db")]
    return {"max_abs_delta_db": float(max(deltas)) if deltas else None, "targets": rows}
