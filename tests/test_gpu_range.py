"""Range compression and raw IQ decode on the GPU against the FP64 oracle
(fast_bp/sar_io.py:18-81, 116-139) and the reference's fractional-delay
known answer (fast_bp/validate.py:55-59)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_fractional_delay_kat_on_gpu(oracle, golden):
    from bp_project_amd.sar_io import range_compress
    ref = golden["synthetic_validation"]
    raw = oracle.lfm_echo(4000, ref["expected_sample"])
    out = range_compress(raw, oversample=4)
    assert out.shape == (1, 16000) and out.dtype == np.complex64
    found = np.argmax(np.abs(out[0])) / 4
    print(f"peak at sample {found} (reference {ref['lfm_peak_sample']})")
    assert found == ref["lfm_peak_sample"]
    exp = oracle.range_compress(raw, oversample=4)
    err = np.linalg.norm(out - exp) / np.linalg.norm(exp)
    assert err <= 1e-5, err


# oversample 1 runs the one-kernel LDS FFT path for transform lengths up to
# 9408 points -- kernels compiled for the length: 1500 -> 2048 (radices
# 8 8 8 4), 2048 -> 2304 (8 8 4 3 3), 600 -> 1024 (8 8 4 4), 2100 -> 2560
# (8 8 8 5); the runtime-plan kernel: 100 -> 512 (8 8 8), 200 -> 640
# (8 4 4 5) -- and hipFFT above (9000 -> 10240); oversampling always takes
# the hipFFT path.
@pytest.mark.parametrize("pulses,samples,oversample,sign,tp", [
    (7, 1500, 1, -1, 1e-6), (5, 2048, 1, 1, 1e-6), (3, 600, 1, 1, 1e-6),
    (3, 100, 1, -1, 1e-6), (2, 200, 1, 1, 1e-6),
    (2, 2100, 1, -1, 1e-6), (2, 9000, 1, -1, 1e-6), (3, 1777, 4, -1, 1e-6),
    (4, 1000, 2, 1, 0.5e-6), (2, 999, 3, -1, 0.25e-6)])
def test_range_compress_vs_oracle(oracle, pulses, samples, oversample, sign, tp):
    from bp_project_amd.sar_io import fft_length, range_compress
    rng = np.random.default_rng(pulses * 1000 + samples)
    raw = (rng.standard_normal((pulses, samples)) + 1j * rng.standard_normal((pulses, samples))
           + 0.3 + 0.2j).astype(np.complex64)
    for p in range(pulses):  # echoes on top of the noise
        raw[p] += 5 * oracle.lfm_echo(samples, 200 + 37.3 * p, tp=tp)[0] * (1 if sign < 0 else 1)
    out = range_compress(raw, tp=tp, bandwidth=200e6, chirp_sign=sign, oversample=oversample)
    exp = oracle.range_compress(raw, tp=tp, bandwidth=200e6, chirp_sign=sign,
                                oversample=oversample, nfft=fft_length(samples, tp))
    assert out.shape == exp.shape == (pulses, samples * oversample)
    err = np.linalg.norm(out - exp) / np.linalg.norm(exp)
    print(f"{pulses}x{samples} os{oversample} sign{sign}: rel err {err:.2e}")
    assert err <= 1e-5


def test_oversample1_transform_length_is_internal():
    """With oversample 1 the library picks its own transform length: any
    valid nfft gives the same output bit for bit."""
    from bp_project_amd.device import DeviceArray
    from bp_project_amd.sar_io import fft_length, range_compress_device
    rng = np.random.default_rng(3)
    raw = (rng.standard_normal((3, 1500)) + 1j * rng.standard_normal((3, 1500))).astype(np.complex64)
    d = DeviceArray.from_host(raw)
    n = fft_length(1500, 1e-6)
    a = range_compress_device(d, nfft=n).get()
    b = range_compress_device(d, nfft=2 * n + 1).get()
    assert np.array_equal(a, b)


def test_decode_frames_vs_oracle(tmp_path, oracle):
    from bp_project_amd import _lib
    from bp_project_amd.device import DeviceArray
    from bp_project_amd.sar_io import Capture
    from capture_files import write_capture
    path = tmp_path / "c.dat"
    iq, _ = write_capture(path, nr=128, frames=20, bad=(0, 7), prefix=11)
    cap = Capture(path)
    sel = [0, 3, 4, 17, 2]
    got = cap.read(sel)
    exp = oracle.decode_frames(path, cap.offset, cap.frame_size, cap.nr, cap.valid_indices[sel])
    assert np.array_equal(got, exp)
    # an index outside the capture decodes to zeros
    frames = cap._frames_on_device()
    idx = DeviceArray.from_host(np.array([5, 999, -1], np.int64), np.int64)
    out = DeviceArray.empty((3, cap.nr))
    _lib.call("bpmi_decode_iq16", _lib.c_void_p(frames.ptr), len(cap.frames), cap.frame_size, 160,
              cap.nr, _lib.c_void_p(idx.ptr), 3, _lib.c_void_p(out.ptr), None)
    o = out.get()
    assert np.array_equal(o[0], oracle.decode_frames(path, cap.offset, cap.frame_size, cap.nr, [5])[0])
    assert not o[1].any() and not o[2].any()
