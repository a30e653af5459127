"""Raw capture frames and range compression (fast_bp/sar_io.py), with the
decode and the matched filter on the GPU.

* ``Capture(path)`` parses a raw ``.dat`` capture like the reference
  (fast_bp/sar_io.py:18-81): the 8-byte frame marker within the first 10 MB,
  the range-sample count from the ``window`` field (offset 60), 192 + 4 nr
  byte frames (160-byte header, interleaved int16 I/Q, 32-byte trailer),
  the header fields at the reference's byte offsets, valid frames by marker,
  navigation scaled to SI units, the constant-parameter checks and the
  summary.  ``read(indices)`` gathers the I/Q blocks of valid frames into
  complex64 rows with ``bpmi_decode_iq16`` (the frames are copied to the
  device once); ``read_device`` keeps the result there.
* ``trajectory`` is the reference's host-side ENU -> (cross-track left,
  along-track, up) conversion (sar_io.py:84-113).
* ``range_compress`` has the reference's signature and semantics
  (sar_io.py:116-139) and runs ``bpmi_range_compress``: mean removal, FFT
  correlation with the centred Hamming-windowed LFM reference, optional
  Fourier oversampling, crop so that sample k is the pulse centred at raw
  sample k.  ``range_compress_device`` takes and returns device arrays.
  Without oversampling the FFTs run at ``transform_length`` rather than the
  reference's next_fast_len: the same linear correlation, to FP32 rounding.

Constants: C, FC = 1.4 GHz, FS = 400 MHz and the frame marker as in
fast_bp/sar_io.py:12-15.
"""
from __future__ import annotations

import ctypes
import json
from pathlib import Path

import numpy as np

from . import _lib
from .device import DeviceArray, stream_handle

C = 299792458.0
FC = 1.4e9
FS = 400e6
MAGIC = bytes([0, 52, 0, 52, 243, 250, 243, 250])
HEADER_BYTES = 160
TRAILER_BYTES = 32
_vp = ctypes.c_void_p

# (name, numpy type, shape, byte offset from the frame marker) -- the frame
# layout the reference reads (fast_bp/sar_io.py:30-37, after read_rawdata_1.m)
_FIELDS = [
    ("magic", "u1", (8,), 0), ("number", "<u4", (), 36), ("bandwidth", "<u4", (), 40),
    ("tp", "<i2", (), 48), ("delay", "<i2", (), 50), ("prt", "<u4", (), 52),
    ("window", "<i2", (), 60), ("heading", "<i2", (), 71),
    ("pitch", "<i2", (), 73), ("roll", "<i2", (), 75),
    ("lat", "<i8", (), 101), ("lon", "<i8", (), 109), ("alt", "<i4", (), 117),
    ("ve", "<i4", (), 121), ("vn", "<i4", (), 125), ("vu", "<i4", (), 129),
    ("lat_gnss", "<i8", (), 133), ("lon_gnss", "<i8", (), 141),
    ("gps_ms", "<u4", (), 149), ("sv", "u1", (), 155),
]
_NAV_SCALE = {"lat": 1e-9, "lon": 1e-9, "lat_gnss": 1e-9, "lon_gnss": 1e-9, "alt": 1e-3,
              "ve": .01, "vn": .01, "vu": .01, "heading": .01, "pitch": .01, "roll": .01}


def frame_dtype(nr):
    """Structured dtype of one frame with ``nr`` range samples."""
    fields = _FIELDS + [("iq", "<i2", (nr, 2), HEADER_BYTES),
                        ("trailer", "u1", (TRAILER_BYTES,), HEADER_BYTES + 4 * nr)]
    return np.dtype({"names": [f[0] for f in fields],
                     "formats": [(f[1], f[2]) if f[2] else f[1] for f in fields],
                     "offsets": [f[3] for f in fields],
                     "itemsize": HEADER_BYTES + 4 * nr + TRAILER_BYTES})


class Capture:
    """Raw capture reader (fast_bp/sar_io.py:18-81)."""

    def __init__(self, path):
        self.path = Path(path)
        with self.path.open("rb") as f:
            prefix = f.read(10_000_000)
        self.offset = prefix.find(MAGIC)
        if self.offset < 0:
            raise ValueError("Frame marker not found in first 10 MB")
        window = int(np.frombuffer(prefix, "<i2", 1, self.offset + 60)[0])
        self.nr = round(window * 1e-8 * FS)
        self.frame_size = HEADER_BYTES + TRAILER_BYTES + self.nr * 4
        size = self.path.stat().st_size
        count = (size - self.offset) // self.frame_size
        all_frames = np.memmap(self.path, dtype=frame_dtype(self.nr), mode="r",
                               offset=self.offset, shape=(count,))
        good = np.all(all_frames["magic"] == np.frombuffer(MAGIC, "u1"), axis=1)
        self.valid_indices = np.flatnonzero(good)
        if not len(self.valid_indices):
            raise ValueError("No valid frames")
        self.frames = all_frames
        self.n = len(self.valid_indices)
        self.nav = {}
        for k in ["number", "gps_ms", "sv", *_NAV_SCALE]:
            self.nav[k] = np.asarray(all_frames[k][good], dtype=np.float64) * _NAV_SCALE.get(k, 1)
        for k in ["tp", "delay", "prt", "bandwidth", "window"]:
            if len(np.unique(np.asarray(all_frames[k][good]))) != 1:
                raise ValueError(f"Changing {k} requires per-pulse handling")
        f = all_frames[self.valid_indices[0]]
        self.tp = float(f["tp"]) * 1e-8
        self.delay = float(f["delay"]) * 1e-8 - .2e-6
        self.prf = 1 / (float(f["prt"]) * 1e-8)
        self.bandwidth = float(f["bandwidth"])
        self.time = (self.nav["number"] - self.nav["number"][0]) / self.prf
        if np.any(np.diff(self.time) <= 0):
            raise ValueError("Non-monotonic pulse numbering")
        self.summary = dict(
            file=str(self.path.resolve()), bytes=size, first_frame_offset=self.offset,
            frame_bytes=self.frame_size, complete_frame_slots=int(count), valid_frames=self.n,
            invalid_slots=int((~good).sum()), first_pulse=int(self.nav["number"][0]),
            last_pulse=int(self.nav["number"][-1]),
            missing_pulses=int(np.sum(np.diff(self.nav["number"]) - 1)),
            range_samples=self.nr, sample_rate_hz=FS, carrier_hz=FC,
            bandwidth_hz=self.bandwidth, pulse_width_s=self.tp, sampling_delay_s=self.delay,
            prf_hz=self.prf, duration_s=float(self.time[-1]),
            discarded_partial_bytes=int((size - self.offset) % self.frame_size))
        self._device_frames = None

    def _frames_on_device(self):
        """Every complete frame slot, as bytes, copied to the device once."""
        if self._device_frames is None:
            raw = np.memmap(self.path, dtype=np.uint8, mode="r", offset=self.offset,
                            shape=(len(self.frames) * self.frame_size,))
            self._device_frames = DeviceArray.from_host(raw, np.uint8)
        return self._device_frames

    def read_device(self, indices, stream=None):
        """complex64 [len(indices), nr] device array of valid frames ``indices``."""
        idx = self.valid_indices[np.asarray(indices)].astype(np.int64).ravel()
        frames = self._frames_on_device()
        d_idx = DeviceArray.from_host(idx, np.int64)
        out = DeviceArray.empty((len(idx), self.nr))
        _lib.call("bpmi_decode_iq16", _vp(frames.ptr), len(self.frames), self.frame_size,
                  HEADER_BYTES, self.nr, _vp(d_idx.ptr), len(idx), _vp(out.ptr),
                  stream_handle(stream))
        return out

    def read(self, indices):
        """complex64 [len(indices), nr] of valid frames ``indices`` (host)."""
        return self.read_device(indices).get()


def trajectory(capture, height=18.0, mode="velocity"):
    """ENU -> cross-track left / along-track / up, scene flat at z = 0
    (fast_bp/sar_io.py:84-113).  Returns (positions [n, 3], geometry dict)."""
    from scipy import integrate
    n, t = capture.nav, capture.time
    velocity = np.column_stack([n["ve"], n["vn"], n["vu"]])
    enu = integrate.cumulative_trapezoid(velocity, t, axis=0, initial=0)
    if mode == "position":
        idx = np.r_[0, np.flatnonzero(np.diff(n["gps_ms"])) + 1]
        lat0 = np.deg2rad(n["lat"][0])
        earth_a = 6378137.
        e2 = 6.69437999014e-3
        rn = earth_a / np.sqrt(1 - e2 * np.sin(lat0) ** 2)
        rm = earth_a * (1 - e2) / (1 - e2 * np.sin(lat0) ** 2) ** 1.5
        ll = np.column_stack([np.deg2rad(n["lon"] - n["lon"][0]) * rn * np.cos(lat0),
                              np.deg2rad(n["lat"] - n["lat"][0]) * rm, n["alt"] - n["alt"][0]])
        enu = np.column_stack([np.interp(t, t[idx], ll[idx, k]) for k in range(3)])
    elif mode == "straight":
        enu = t[:, None] * velocity.mean(axis=0)
    along = velocity[:, :2].mean(axis=0)
    along /= np.linalg.norm(along)
    left = np.array([-along[1], along[0]])
    xy = enu[:, :2] - enu[len(t) // 2, :2]
    pos = np.column_stack([xy @ left, xy @ along, enu[:, 2] + height])
    return pos, dict(mode=mode, height_reference_m=height, along_enu=along.tolist(),
                     left_enu=left.tolist(),
                     mean_speed_m_s=float(np.linalg.norm(velocity[:, :2], axis=1).mean()),
                     mean_course_deg=float(np.rad2deg(np.arctan2(along[0], along[1])) % 360),
                     extent_m=np.ptp(pos, axis=0).tolist())


def fft_length(samples, tp, fs=FS):
    """nfft of the reference: scipy.fft.next_fast_len(samples + L - 1)."""
    from scipy import fft
    return int(fft.next_fast_len(int(samples) + int(round(tp * fs)) - 1))


def transform_length(samples, tp, oversample=1, fs=FS, nfft=None):
    """FFT length bpmi_range_compress runs: with oversample 1 the smallest
    2^a m (m = 1, 3, 5, 9) >= samples + L - 1 (the same linear correlation,
    faster in rocFFT), else nfft (default: the reference's fft_length)."""
    nfft = fft_length(samples, tp, fs) if nfft is None else int(nfft)
    if oversample != 1:
        return nfft
    need = int(samples) + int(round(tp * fs)) - 1
    best = None
    for m in (1, 3, 5, 9):
        n = m
        while n < need:
            n *= 2
        best = n if best is None else min(best, n)
    return best


def range_compress_device(raw, tp=1e-6, bandwidth=200e6, chirp_sign=-1, oversample=1, *,
                          fs=FS, nfft=None, out=None, stream=None):
    """``range_compress`` on a device array raw complex64 [pulses, samples];
    returns the device array complex64 [pulses, samples * oversample]."""
    if raw.dtype != np.complex64 or raw.ndim != 2:
        raise TypeError("raw data must be a complex64 [pulse, sample] device array")
    pulses, samples = raw.shape
    nfft = fft_length(samples, tp, fs) if nfft is None else int(nfft)
    nbytes = ctypes.c_size_t()
    _lib.call("bpmi_range_compress_workspace_bytes", pulses, samples, float(tp), float(fs), nfft,
              int(oversample), ctypes.byref(nbytes))
    work = DeviceArray.empty(-(-nbytes.value // 8))
    if out is None:
        out = DeviceArray.empty((pulses, samples * int(oversample)))
    _lib.call("bpmi_range_compress", _vp(raw.ptr), pulses, samples, samples, float(tp),
              float(fs), float(bandwidth), int(chirp_sign), nfft, int(oversample), _vp(out.ptr),
              _vp(work.ptr), work.nbytes, stream_handle(stream))
    return out


def range_compress(raw, tp=1e-6, bandwidth=200e6, chirp_sign=-1, oversample=1):
    """fast_bp/sar_io.py:116-139 with the same signature: host complex array
    [pulses, samples] in, complex64 [pulses, samples * oversample] out."""
    raw = np.ascontiguousarray(raw, dtype=np.complex64)
    if raw.ndim != 2:
        raise ValueError("raw must be [pulse, sample]")
    out = range_compress_device(DeviceArray.from_host(raw), tp, bandwidth, chirp_sign, oversample)
    return out.get()


def write_json(path, content):
    """fast_bp/sar_io.py:142-143."""
    Path(path).write_text(json.dumps(content, indent=2, ensure_ascii=False), encoding="utf-8")
