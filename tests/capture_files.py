"""Synthetic raw capture files in the frame layout that fast_bp/sar_io.py:18-81
reads (test data writer; the reference's own capture is not in its
repository).  Each frame: 8-byte marker, header fields at their byte offsets,
interleaved int16 I/Q from byte 160, a 32-byte trailer."""
import numpy as np

MAGIC = bytes([0, 52, 0, 52, 243, 250, 243, 250])


def write_capture(path, nr=64, frames=12, prefix=37, bad=(3,), tail=5, seed=0,
                  tp=100, delay=500, prt=100000, bandwidth=200_000_000, vary=None):
    """Write ``frames`` frames of ``nr`` samples after ``prefix`` junk bytes;
    frames listed in ``bad`` get a broken marker; ``tail`` partial bytes at
    the end.  Returns (iq int16 [frames, nr, 2], header values per frame)."""
    rng = np.random.default_rng(seed)
    assert nr % 4 == 0
    window = nr // 4  # nr = round(window * 1e-8 * 400e6)
    size = 192 + 4 * nr
    iq = rng.integers(-2000, 2000, size=(frames, nr, 2)).astype("<i2")
    header = []
    blob = bytearray(rng.integers(0, 255, prefix, dtype=np.uint8).tobytes().replace(MAGIC[:1], b"\x01"))
    for i in range(frames):
        f = bytearray(size)
        f[0:8] = MAGIC if i not in bad else b"\x00" * 8
        h = dict(number=1000 + 2 * i + (1 if i >= 6 else 0), bandwidth=bandwidth, tp=tp,
                 delay=delay, prt=prt, window=window, heading=1234 + i, pitch=-5, roll=7,
                 lat=31_123_456_789 + 10 * i, lon=121_987_654_321 - 10 * i, alt=45_678 + i,
                 ve=150 + i, vn=-2000 - i, vu=3, lat_gnss=31_123_456_000, lon_gnss=121_987_654_000,
                 gps_ms=10 * (i // 3), sv=9)
        if vary and i == frames - 1:
            h[vary] = h[vary] + 1
        for name, fmt, off in (("number", "<u4", 36), ("bandwidth", "<u4", 40), ("tp", "<i2", 48),
                               ("delay", "<i2", 50), ("prt", "<u4", 52), ("window", "<i2", 60),
                               ("heading", "<i2", 71), ("pitch", "<i2", 73), ("roll", "<i2", 75),
                               ("lat", "<i8", 101), ("lon", "<i8", 109), ("alt", "<i4", 117),
                               ("ve", "<i4", 121), ("vn", "<i4", 125), ("vu", "<i4", 129),
                               ("lat_gnss", "<i8", 133), ("lon_gnss", "<i8", 141),
                               ("gps_ms", "<u4", 149), ("sv", "u1", 155)):
            b = np.array([h[name]], dtype=fmt).tobytes()
            f[off:off + len(b)] = b
        f[160:160 + 4 * nr] = iq[i].tobytes()
        f[160 + 4 * nr:] = bytes(range(32))
        blob += f
        header.append(h)
    blob += b"\x07" * tail
    with open(path, "wb") as fh:
        fh.write(bytes(blob))
    return iq, header
