"""Seeded synthetic inputs: the reference's validation scenes and the
stripmap point-target workloads this implementation is measured on.

All generators return plain numpy arrays (complex64 range-compressed data
``[pulse, bin]``, FP64 positions ``[pulse, 3]``, FP64 weight and phase per
pulse) plus the scalars the BP entry points take.  Nothing here touches the
GPU.

Reference scenes (exact reproductions of the reference's seeded inputs):
* ``operator_validation_cases`` -- fast_bp/gpu_baseline.py:125-160
* ``resident_validation_cases`` -- fast_bp/resident_experiment.py:146-156
  (seed 20260917) and fast_bp/resident_followup.py:48-58 (seed 20260918)
* ``synthetic_kat_scene``       -- fast_bp/validate.py:13-31

Stripmap workloads (x = cross-track, y = along-track, z = altitude):
straight track at 3000 m altitude, 100 m/s, PRF 500 Hz (0.2 m pulse
spacing), FP64 positions with seeded Gaussian jitter; X-band LFM with
fc = 9.6 GHz, B = 150 MHz, fs = 180 MHz (bin spacing c/(2 fs)); point targets
with amplitude U(0.5, 1) and uniform phase, range-compressed analytically as
``a exp(-j k d) sinc((r - d) / (c / 2B))`` (sinc truncated to +-48 bins around
each target), plus complex Gaussian clutter at 20 dB SNR.
"""
from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np
from scipy.signal.windows import hann

C = 299792458.0
FC_REF = 1.4e9          # reference carrier (fast_bp/sar_io.py:13)
FS_REF = 400e6          # reference sample rate (fast_bp/sar_io.py:14)
SINC_HALF_BINS = 48


def _validation_geometry(pulses, wobble_y):
    slow = np.arange(pulses, dtype=np.float64)
    y = np.linspace(-15.0, 15.0, pulses)
    if wobble_y:
        y = y + 0.03 * np.sin(slow)
    return np.column_stack([0.05 * np.sin(0.37 * slow), y, 18.0 + 0.1 * np.cos(0.23 * slow)])


def operator_validation_cases(seed=20260916):
    """Seeded operator validation of fast_bp/gpu_baseline.py:121-189.

    Yields dicts in the reference's order ((k sign, side) = (-1, 1), (1, -1)),
    drawing the Gaussian data from one generator in that order.
    """
    rng = np.random.default_rng(seed)
    for sign, side in ((-1, 1), (1, -1)):
        pulses, samples = 31, 512
        position = _validation_geometry(pulses, wobble_y=True)
        data = (rng.standard_normal((pulses, samples))
                + 1j * rng.standard_normal((pulses, samples))).astype(np.complex64)
        fc = sign * FC_REF
        yield dict(sign=sign, side=side, data=data, pos=position, weight=hann(pulses),
                   phase=np.linspace(-0.2, 0.3, pulses), r_start=900.0, data_dr=0.5, fc=fc,
                   k=4.0 * np.pi * fc / C, extent=(1080.0, 1120.0, -0.05, 0.05),
                   leaf_size=8, range_spacing=0.5, angular_oversample=3.0)


def resident_validation_cases(seed=20260917):
    """Seeded resident/batched validations (resident_experiment.py:146-156,
    resident_followup.py:48-58): four (sign, side) cases from one generator."""
    rng = np.random.default_rng(seed)
    k_abs = 4 * np.pi * FC_REF / C
    for sign, side in ((-1, 1), (-1, -1), (1, 1), (1, -1)):
        n, ns = 31, 512
        position = _validation_geometry(n, wobble_y=False)
        data = (rng.normal(size=(n, ns)) + 1j * rng.normal(size=(n, ns))).astype(np.complex64)
        yield dict(sign=sign, side=side, data=data, pos=position, weight=hann(n),
                   phase=np.linspace(-0.2, 0.3, n), r_start=900.0, data_dr=0.5,
                   fc=sign * FC_REF, k=sign * k_abs, extent=(1080.0, 1120.0, -0.05, 0.05),
                   leaf_size=8, range_spacing=0.5, angular_oversample=3.0)


def synthetic_kat_scene():
    """Deterministic analytic scene of fast_bp/validate.py:13-31 (no RNG)."""
    k = 4 * np.pi * FC_REF / C
    n = 512
    t = np.linspace(-1, 1, n)
    pos = np.column_stack([.08 * np.sin(5 * t), 50 * t + .4 * t * t, 18 + .15 * np.sin(2 * t)])
    targets = np.array([[710.31, -145.27, 0.], [987.21, 31.72, 0.], [1190.77, -380.11, 0.]])
    strength = np.array([1, complex(.7, .3), complex(.5, -.2)])
    r_start = 500.
    ddr = C / (2 * FS_REF)
    rr = r_start + ddr * np.arange(2300)
    data = np.zeros((n, len(rr)), np.complex64)
    for xyz, a in zip(targets, strength):
        distance = np.linalg.norm(pos - xyz, axis=1)
        data += (a * np.exp(-1j * k * distance)[:, None]
                 * np.sinc((rr[None, :] - distance[:, None]) / (C / (2 * 200e6)))).astype(np.complex64)
    w = hann(n)
    phase = np.zeros(n)
    cuts = []
    for xyz in targets:
        x = xyz[0] + np.arange(-3, 3.01, .15)
        y = xyz[1] + np.arange(-8, 8.01, .2)
        xx, yy = np.meshgrid(x, y)
        cuts.append(np.column_stack([xx.ravel(), yy.ravel(), np.zeros(xx.size)]))
    return dict(data=data, pos=pos, weight=w, phase=phase, r_start=r_start, data_dr=ddr,
                fc=FC_REF, k=k, targets=targets, strength=strength, cuts=cuts,
                points=np.vstack(cuts), extent=(650, 1300, -.5, .18), range_spacing=.25,
                angular_oversample=3.)


def autofocus_kat_scene():
    """Deterministic phase-centre error scene of fast_bp/validate_autofocus.py:9-20
    (no RNG): 4096 pulses along a straight 140 m track at 18 m height, four
    targets, range-compressed with a 400 MHz sinc (900 bins from 650 m), and a
    known detrended sinusoidal x/y position error ``err`` that the phase-gradient
    update must recover from data recorded on ``truth`` but imaged on
    ``truth + err``."""
    from scipy import signal
    n = 4096
    t = np.linspace(0, 4, n)
    truth = np.column_stack([np.zeros(n), np.linspace(-70, 70, n), np.full(n, 18.)])
    targets = np.array([[710., -290., 0.], [720., 320., 0.], [810., -100., 0.], [870., 280., 0.]])
    k = -4 * np.pi * FC_REF / C
    err = np.column_stack([.06 * np.sin(2 * np.pi * t), .12 * np.cos(np.pi * t), np.zeros(n)])
    err[:, :2] = signal.detrend(err[:, :2], axis=0)
    r0 = 650.
    dr = C / (2 * FS_REF)
    rr = r0 + dr * np.arange(900)
    data = np.zeros((n, len(rr)), np.complex64)
    for xyz in targets:
        d = np.linalg.norm(truth - xyz, axis=1)
        data += (np.exp(-1j * k * d)[:, None]
                 * np.sinc((rr[None, :] - d[:, None]) / (C / 400e6))).astype(np.complex64)
    return dict(data=data, truth=truth, err=err, pos=truth + err, targets=targets,
                r_start=r0, data_dr=dr, k=k, prf=1024., cutoff=15.)


# ---------------------------------------------------------------------------
# Stripmap point-target workloads
# ---------------------------------------------------------------------------

FC_X = 9.6e9
BANDWIDTH = 150e6
FS_X = 180e6
ALTITUDE = 3000.0
SPEED = 100.0
PRF = 500.0


@dataclass
class CartesianGrid:
    """Regular ground grid on z = z0; image layout [ny, nx] (x fastest),
    like the reference's Cartesian ground image (fast_bp/run.py:118-125)."""
    x0: float
    dx: float
    nx: int
    y0: float
    dy: float
    ny: int
    z0: float = 0.0

    @property
    def shape(self):
        return self.ny, self.nx

    @property
    def x(self):
        return self.x0 + self.dx * np.arange(self.nx)

    @property
    def y(self):
        return self.y0 + self.dy * np.arange(self.ny)

    def points(self):
        yy, xx = np.meshgrid(self.y, self.x, indexing="ij")
        return np.column_stack([xx.ravel(), yy.ravel(), np.full(xx.size, self.z0)])


@dataclass
class Scene:
    data: np.ndarray
    pos: np.ndarray
    weight: np.ndarray
    phase: np.ndarray
    r_start: float
    data_dr: float
    fc: float
    k: float
    grid: CartesianGrid
    targets: np.ndarray
    amplitudes: np.ndarray
    true_pos: np.ndarray
    meta: dict = field(default_factory=dict)

    @property
    def pulses(self):
        return self.data.shape[0]

    @property
    def bins(self):
        return self.data.shape[1]


SCENE_ROWS = 512     # pulses per generation chunk (fixes the clutter streams)


def _scene_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _fill_rows(data, p0, p1, true_pos, targets, amp, k, r_start, data_dr, phase_error, scale,
               seed):
    """Rows [p0, p1) of a stripmap scene: every target's truncated sinc
    response (targets in order), the per-pulse phase error, then complex
    Gaussian clutter drawn from the chunk's own seeded stream, so the scene
    does not depend on how many threads fill it."""
    rows = data[p0:p1]
    rows[...] = 0
    bins = data.shape[1]
    res = C / (2 * BANDWIDTH)
    offsets = np.arange(-SINC_HALF_BINS, SINC_HALF_BINS + 1)
    flat = rows.reshape(-1)
    base = (np.arange(p1 - p0) * bins)[:, None]
    for t in range(len(targets)):
        d = np.linalg.norm(true_pos[p0:p1] - targets[t], axis=1)
        centre = np.rint((d - r_start) / data_dr).astype(np.int64)
        idx = centre[:, None] + offsets[None, :]
        valid = (idx >= 0) & (idx < bins)
        vals = amp[t] * np.exp(-1j * k * d)[:, None] * np.sinc(
            (r_start + data_dr * idx - d[:, None]) / res)
        # (row, bin) pairs are distinct within one target, so += is exact
        flat[(base + idx)[valid]] += vals[valid].astype(np.complex64)
    if np.any(phase_error[p0:p1]):
        rows *= np.exp(1j * phase_error[p0:p1]).astype(np.complex64)[:, None]
    if scale is not None:
        g = np.random.default_rng([seed, p0 // SCENE_ROWS, 0x5CE7E])
        rows.real += g.standard_normal(rows.shape, dtype=np.float32) * scale
        rows.imag += g.standard_normal(rows.shape, dtype=np.float32) * scale


def stripmap_scene(pulses, bins, n_targets, grid_n, grid_spacing, seed, *, ground_range=5000.0,
                   slant_range=None, patch=None, curved=False, motion_rms=0.0,
                   phase_walk_rms=0.0, along_track_error=0.0, snr_db=20.0, target_margin=0.1,
                   threads=None):
    """Seeded stripmap scene.

    ``patch`` is the side of the square target patch in metres (default: the
    output grid extent).  ``slant_range`` overrides ``ground_range`` for the
    scene centre.  ``curved`` bends the nominal track (quadratic cross-track
    bow of 10 m plus a 2 m altitude swell over the aperture); ``motion_rms``
    adds a seeded 3-axis random walk scaled to that RMS (metres) to the
    antenna path, which the positions ``pos`` include (a known, irregular
    track).  ``phase_walk_rms`` (rad) and ``along_track_error`` (m)
    corrupt the data relative to the nominal ``pos`` for autofocus studies.
    The data are filled in chunks of SCENE_ROWS pulses on ``threads`` threads
    (default: the usable cores, at most 16); the clutter of each chunk comes
    from its own stream seeded by (seed, chunk), so the scene is the same for
    any thread count.
    """
    rng = np.random.default_rng(seed)
    if slant_range is not None:
        ground_range = float(np.sqrt(slant_range ** 2 - ALTITUDE ** 2))
    k = 4 * np.pi * FC_X / C
    data_dr = C / (2 * FS_X)
    spacing = SPEED / PRF
    y = (np.arange(pulses) - (pulses - 1) / 2) * spacing
    nominal = np.column_stack([np.zeros(pulses), y, np.full(pulses, ALTITUDE)])
    if curved:
        s = y / max(1.0, np.abs(y).max())
        nominal[:, 0] += 10.0 * s * s
        nominal[:, 2] += 2.0 * np.sin(np.pi * s)
    pos = nominal + rng.normal(scale=1e-3, size=nominal.shape)
    if motion_rms > 0:
        walk = np.cumsum(rng.standard_normal((pulses, 3)), axis=0)
        walk -= walk.mean(axis=0)
        walk *= motion_rms / np.sqrt(np.mean(walk ** 2, axis=0))
        pos += walk
    true_pos = pos.copy()
    if along_track_error > 0:
        drift = np.cumsum(rng.standard_normal(pulses))
        drift -= drift.mean()
        drift *= along_track_error / np.sqrt(np.mean(drift ** 2))
        true_pos[:, 1] += drift
    extent = grid_n * grid_spacing
    patch = extent if patch is None else float(patch)
    half = 0.5 * patch * (1 - target_margin)
    tx = ground_range + rng.uniform(-half, half, n_targets)
    ty = rng.uniform(-half, half, n_targets)
    targets = np.column_stack([tx, ty, np.zeros(n_targets)])
    amp = rng.uniform(0.5, 1.0, n_targets) * np.exp(1j * rng.uniform(0, 2 * np.pi, n_targets))
    centre_slant = float(np.hypot(ground_range, ALTITUDE))
    r_start = centre_slant - 0.5 * bins * data_dr
    phase_error = np.zeros(pulses)
    if phase_walk_rms > 0:
        walk = np.cumsum(rng.standard_normal(pulses))
        walk -= walk.mean()
        phase_error = walk * phase_walk_rms / np.sqrt(np.mean(walk ** 2))
    data = np.empty((pulses, bins), np.complex64)
    scale = None
    if snr_db is not None:
        scale = np.float32(np.sqrt((7.0 / 12.0) / 10 ** (snr_db / 10) / 2))
    chunks = range(0, pulses, SCENE_ROWS)
    with ThreadPoolExecutor(max_workers=threads or _scene_threads()) as pool:
        list(pool.map(lambda p0: _fill_rows(data, p0, min(pulses, p0 + SCENE_ROWS), true_pos,
                                            targets, amp, k, r_start, data_dr, phase_error,
                                            scale, seed), chunks))
    g0 = grid_n * grid_spacing / 2
    grid = CartesianGrid(x0=ground_range - g0, dx=grid_spacing, nx=grid_n, y0=-g0,
                         dy=grid_spacing, ny=grid_n)
    return Scene(data=data, pos=pos, weight=np.ones(pulses), phase=np.zeros(pulses),
                 r_start=r_start, data_dr=data_dr, fc=FC_X, k=k, grid=grid, targets=targets,
                 amplitudes=amp, true_pos=true_pos,
                 meta=dict(seed=seed, ground_range=ground_range, centre_slant=centre_slant,
                           phase_error=phase_error, spacing=spacing, snr_db=snr_db))


# Workload definitions: (pulses, bins, targets, grid_n, spacing, extra kwargs).
WORKLOADS = {
    "c1_direct_small": dict(pulses=1024, bins=2048, n_targets=25, grid_n=256, grid_spacing=0.5,
                            seed=20260917, patch=128.0),
    "c2_ffbp_full": dict(pulses=8192, bins=8192, n_targets=200, grid_n=4096, grid_spacing=0.25,
                         seed=20260918),
    "c3_direct_precision": dict(pulses=16384, bins=16384, n_targets=200, grid_n=4096,
                                grid_spacing=0.25, seed=20260919, slant_range=15000.0,
                                curved=True, motion_rms=0.05),
    "c4_autofocus": dict(pulses=4096, bins=4096, n_targets=50, grid_n=1024, grid_spacing=0.5,
                         seed=20260920, phase_walk_rms=0.5, along_track_error=0.02),
}


def workload_scene(name, **overrides):
    spec = dict(WORKLOADS[name])
    spec.update(overrides)
    return stripmap_scene(**spec)


def polar_extent_for_grid(grid, origin, side=1, margin=2.0):
    """(slant_min, slant_max, u_min, u_max) about ``origin`` covering ``grid``
    (plus ``margin`` metres), the extent argument of the FFBP plan."""
    xs = np.array([grid.x0 - margin, grid.x0 + grid.dx * (grid.nx - 1) + margin])
    ys = np.array([grid.y0 - margin, grid.y0 + grid.dy * (grid.ny - 1) + margin])
    xx, yy = np.meshgrid(np.linspace(xs[0], xs[1], 65), np.linspace(ys[0], ys[1], 65))
    dx = side * (xx - origin[0])
    dy = yy - origin[1]
    g = np.hypot(dx, dy)
    r = np.sqrt(g * g + (grid.z0 - origin[2]) ** 2)
    u = dy / g
    return float(r.min()), float(r.max()), float(u.min()), float(u.max())
