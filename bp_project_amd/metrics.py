"""Image-quality metrics with the reference's conventions (host numpy code).

* ``fidelity(reference, estimate)`` -- relative complex error (NRMSE),
  SER, coherence, magnitude error, finite fraction and bias-removed phase
  statistics above -30 dB, peak shift (fast_bp/gpu_baseline.py:49-94).
* ``focus_metrics(z)`` -- entropy of normalised intensity, contrasts
  (sar_quality/metrics.py:6-22).
* ``impulse_metrics(x, z)`` -- 1-D half-power IRW, first-minima mainlobe,
  PSLR and ISLR (sar_quality/metrics.py:42-74).
* ``profile_metrics`` -- the qualified-null 1-D point-response measurement
  (the convention of vancouver_bp/quality.py:73-176), ``cut_deltas`` and
  ``pslr_islr_deltas`` -- its PSLR/ISLR differences between two images.
* ``peak_phase_error`` -- phase difference at each target's peak pixel.
* ``timing_summary`` -- median/IQR of timing samples.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.ndimage import gaussian_filter1d
from scipy.signal import find_peaks


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


# ---------------------------------------------------------------------------
# 1-D point-response measurements
#
# A cut z(x) through a point target is measured on its normalised power
# p = |z|^2 / |z(peak)|^2.  The impulse-response width is the distance between
# the two half-power crossings (linear interpolation between samples).  The
# mainlobe ends at the first null on each side; PSLR is the highest power
# outside the mainlobe and ISLR the power integrated (trapezoid rule) outside
# it over the power integrated inside it, both in dB.
#
# How the nulls are found fixes the metric, so the constants below are the
# reference's (vancouver_bp/quality.py:73-176, "qualified first minima"):
# minima are looked for on the power smoothed by a Gaussian of
# NULL_SMOOTH_PER_IRW x IRW samples, clamped to NULL_SMOOTH_LIMITS; a null is
# the minimum nearest the mainlobe whose unsmoothed power is at most
# NULL_DEPTHS[0] of the peak, or failing that NULL_DEPTHS[1]; the ratios are
# reliable when both nulls lie at or below RELIABLE_NULL_DB, the integration
# window is not truncated and the half-power region is one piece.
# impulse_metrics is the plain variant of sar_quality/metrics.py:42-74:
# nulls are the nearest raw local minima.
# ---------------------------------------------------------------------------
NULL_SMOOTH_PER_IRW = 0.04
NULL_SMOOTH_LIMITS = (0.5, 4.0)
NULL_DEPTHS = (0.1, 0.25)
RELIABLE_NULL_DB = -15.0


def _level_crossing(xa, xb, pa, pb, level=0.5):
    """x where the line through (xa, pa), (xb, pb) reaches ``level``."""
    if pa == pb:
        return float(0.5 * (xa + xb))
    return float(xa + (xb - xa) * (level - pa) / (pb - pa))


def _half_power(x, p, peak):
    """(last index left of the peak with p <= 1/2, first such index right of
    it, left crossing, right crossing), or None when a side never drops."""
    left = np.flatnonzero(p[:peak] <= 0.5)
    right = np.flatnonzero(p[peak + 1:] <= 0.5)
    if left.size == 0 or right.size == 0:
        return None
    lo, hi = int(left[-1]), int(right[0]) + peak + 1
    return (lo, hi, _level_crossing(x[lo], x[lo + 1], p[lo], p[lo + 1]),
            _level_crossing(x[hi - 1], x[hi], p[hi - 1], p[hi]))


def _local_minima(v):
    return find_peaks(-v)[0]


def _qualified_nulls(x, p, lo, hi, irw):
    """Indices of the first qualified null left of ``lo`` and right of
    ``hi`` (None where there is none)."""
    step = float(np.median(np.diff(x)))
    sigma = float(np.clip(NULL_SMOOTH_PER_IRW * irw / step, *NULL_SMOOTH_LIMITS))
    mins = _local_minima(gaussian_filter1d(p, sigma, mode="nearest"))
    interior = mins[(mins > 0) & (mins < len(p) - 1)]
    sides = (interior[interior < lo][::-1], interior[interior > hi])
    nulls = []
    for cand in sides:
        pick = None
        for depth in NULL_DEPTHS:
            deep = cand[p[cand] <= depth]
            if deep.size:
                pick = int(deep[0])
                break
        nulls.append(pick)
    return nulls[0], nulls[1]


def _lobe_ratios(x, p, ml, mr):
    """(PSLR dB, ISLR dB) for the mainlobe [ml, mr] of normalised power p."""
    inside = float(np.trapezoid(p[ml:mr + 1], x[ml:mr + 1]))
    outside = float(np.trapezoid(p[:ml + 1], x[:ml + 1]) + np.trapezoid(p[mr:], x[mr:]))
    peak_side = float(max(p[:ml].max(initial=0.0), p[mr + 1:].max(initial=0.0)))
    return _db(peak_side), _db(outside / inside)


def impulse_metrics(x, z):
    """1-D half-power IRW, mainlobe between the nearest raw minima, and the
    PSLR/ISLR of the finite cut."""
    x = np.asarray(x, float)
    p = np.abs(np.asarray(z)) ** 2
    if not np.isfinite(p).all() or p.max() <= 0:
        return dict(valid=False, reason="power is not finite and positive")
    peak = int(np.argmax(p))
    p = p / p[peak]
    hp = _half_power(x, p, peak)
    if hp is None:
        return dict(valid=False, reason="no half-power point on one side of the peak")
    lo, hi, xl, xr = hp
    out = dict(valid=True, irw_m=float(xr - xl), peak_offset_m=float(x[peak]))
    mins = _local_minima(p)
    left, right = mins[mins < lo], mins[mins > hi]
    if left.size == 0 or right.size == 0:
        return dict(out, sidelobes_valid=False, reason="no minimum beyond the half-power points")
    ml, mr = int(left[-1]), int(right[0])
    pslr, islr = _lobe_ratios(x, p, ml, mr)
    return dict(out, sidelobes_valid=True, pslr_db=pslr, islr_db=islr,
                mainlobe_bounds_m=[float(x[ml]), float(x[mr])])


def profile_metrics(x, z, *, expected_irw_m=None, integration_half_width_m=None, peak_index=None):
    """Point response of one cut with the qualified-null convention above.

    ``integration_half_width_m`` restricts the measurement to that distance
    either side of the peak; ``peak_index`` names the peak sample (default:
    the cut's maximum).  Returns a dict with ``valid`` (IRW measured),
    ``sidelobes_valid`` (PSLR/ISLR reliable), ``irw_m``, ``pslr_db``,
    ``islr_db``, the mainlobe bounds and null depths, and ``reason`` or
    ``warnings`` where something is missing or doubtful."""
    x = np.asarray(x, np.float64)
    z = np.asarray(z)
    if x.ndim != 1 or x.shape != z.shape or x.size < 5:
        raise ValueError("x and z must be 1-D, of equal length, with at least 5 samples")
    if not np.isfinite(x).all() or np.any(np.diff(x) <= 0):
        raise ValueError("x must be finite and strictly increasing")
    power = np.abs(z.astype(np.complex128)) ** 2
    out = dict(valid=False, sidelobes_valid=False, warnings=[])
    if not np.isfinite(power).all() or not np.any(power > 0):
        return dict(out, reason="power is not finite and positive")
    peak = int(np.argmax(power)) if peak_index is None else int(peak_index)
    if not (0 <= peak < power.size) or power[peak] <= 0:
        raise ValueError("peak_index must name a sample of the cut with positive power")
    keep = np.arange(x.size)
    truncated = False
    if integration_half_width_m is not None:
        half = float(integration_half_width_m)
        if half <= 0:
            raise ValueError("integration_half_width_m must be positive")
        centre = x[peak]
        truncated = bool(centre - half < x[0] or centre + half > x[-1])
        keep = np.flatnonzero(np.abs(x - centre) <= half)
        x, power = x[keep], power[keep]
    if x.size < 5:
        return dict(out, reason="fewer than 5 samples inside the integration window")
    peak -= int(keep[0])
    p = power / power[peak]
    out.update(peak_coordinate_m=float(x[peak]), peak_sample_index=int(keep[peak]),
               step_m=float(np.median(np.diff(x))), integration_bounds_m=[float(x[0]), float(x[-1])],
               integration_window_truncated=truncated)
    hp = _half_power(x, p, peak)
    if hp is None:
        return dict(out, reason="no half-power point on one side of the peak")
    lo, hi, xl, xr = hp
    out.update(valid=True, irw_m=float(xr - xl), half_power_bounds_m=[xl, xr])
    if expected_irw_m is not None:
        out["irw_to_expected_ratio"] = out["irw_m"] / float(expected_irw_m)
    ml, mr = _qualified_nulls(x, p, lo, hi, out["irw_m"])
    if ml is None or mr is None:
        return dict(out, reason="no qualified null on one side; no sidelobe ratios")
    if ml == 0 or mr == p.size - 1:
        return dict(out, reason="a null lies on the window edge")
    pslr, islr = _lobe_ratios(x, p, ml, mr)
    depths = [_db(p[ml]), _db(p[mr])]
    rises = np.diff(np.r_[0, (p >= 0.5).astype(np.int8), 0])
    pieces = int(np.count_nonzero(rises == 1))
    deep = max(depths) <= RELIABLE_NULL_DB
    out.update(sidelobes_valid=bool(deep and not truncated and pieces == 1), pslr_db=pslr,
               islr_db=islr, mainlobe_bounds_m=[float(x[ml]), float(x[mr])],
               mainlobe_minima_db=depths, mainlobe_sample_indices=[int(keep[ml]), int(keep[mr])],
               half_power_components=pieces)
    if not deep:
        out["warnings"].append(f"a null is shallower than {RELIABLE_NULL_DB:g} dB")
    if truncated:
        out["warnings"].append("integration window extends beyond the cut")
    if pieces != 1:
        out["warnings"].append("the half-power region has more than one piece")
    return out


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


def cut_deltas(axis, ref_cut, est_cut, peak_index, half_width_m=8.0):
    """PSLR/ISLR of one 1-D cut of ``est`` minus the same of ``ref`` (both
    measured about ``peak_index`` within +-half_width_m), or None when either
    has no sidelobe ratios.  A decreasing axis is measured reversed."""
    axis = np.asarray(axis, np.float64)
    ref_cut, est_cut = np.asarray(ref_cut), np.asarray(est_cut)
    if axis.size > 1 and axis[1] < axis[0]:
        axis, ref_cut, est_cut = axis[::-1], ref_cut[::-1], est_cut[::-1]
        peak_index = axis.size - 1 - peak_index
    a = profile_metrics(axis, ref_cut, integration_half_width_m=half_width_m,
                        peak_index=peak_index)
    b = profile_metrics(axis, est_cut, integration_half_width_m=half_width_m,
                        peak_index=peak_index)
    if "pslr_db" not in a or "pslr_db" not in b:
        return None
    return dict(pslr_delta_db=b["pslr_db"] - a["pslr_db"], islr_delta_db=b["islr_db"] - a["islr_db"],
                pslr_ref_db=a["pslr_db"], islr_ref_db=a["islr_db"])


def pslr_islr_deltas(reference, estimate, grid, peaks, half_width_m=8.0):
    """PSLR/ISLR (dB) of the range (x) and azimuth (y) cuts through each peak
    (row, col) of two images [ny, nx] on ``grid``: estimate minus reference.
    Returns the worst absolute delta and the per-target records."""
    rows = []
    for (i, j) in peaks:
        rec = {}
        for name, axis, ref_cut, est_cut, at in (
                ("range", grid.x, reference[i, :], estimate[i, :], j),
                ("azimuth", grid.y, reference[:, j], estimate[:, j], i)):
            d = cut_deltas(axis, ref_cut, est_cut, at, half_width_m)
            if d is not None:
                rec[name] = d
        rows.append(rec)
    worst = [abs(v[key]) for rec in rows for v in rec.values()
             for key in ("pslr_delta_db", "islr_delta_db")]
    return {"max_abs_delta_db": float(max(worst)) if worst else None, "targets": rows}
