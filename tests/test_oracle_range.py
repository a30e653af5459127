"""Range-compression oracle (FP64 restatement of fast_bp/sar_io.py:116-139)
against the reference's fractional-delay known answer and closed forms."""
import numpy as np


def test_fractional_delay_kat(oracle, golden):
    """fast_bp/validate.py:55-59: 4000 samples, chirp centred at 1512.4,
    4x oversampling; the reference found the peak at sample 1512.5."""
    ref = golden["synthetic_validation"]
    raw = oracle.lfm_echo(4000, ref["expected_sample"])
    out = oracle.range_compress(raw, oversample=4)
    assert out.shape == (1, 16000)
    found = np.argmax(np.abs(out[0])) / 4
    assert found == ref["lfm_peak_sample"] == 1512.5
    assert abs(found - ref["expected_sample"]) <= .25


def test_integer_delay_peak_value_and_length(oracle):
    """A chirp centred on integer sample k compresses to a peak at k whose
    value is the windowed energy sum_n w_n (|x_n| = 1 under the window) minus
    the mean term; the output keeps the raw length."""
    tp, B, fs = 1e-6, 200e6, 400e6
    L = int(round(tp * fs))
    for k in (700, 1000):
        raw = oracle.lfm_echo(3000, k)
        out = oracle.range_compress(raw)
        assert out.shape == raw.shape
        assert np.argmax(np.abs(out[0])) == k
        ref = oracle.lfm_reference(tp, B, -1, fs)
        x = raw[0].astype(np.complex128)
        x = x - x.mean()
        # correlation at lag 0 around the centre: sum_n x[k - L//2 + n] conj(ref[n])
        direct = np.sum(x[k - L // 2:k - L // 2 + L] * np.conj(ref))
        assert abs(out[0, k] - direct) <= 1e-9 * abs(direct)
        assert abs(abs(direct) - np.sum(np.hamming(L + 1)[:-1])) < 0.02 * abs(direct)


def test_linear_in_the_data_and_shift_equivariant(oracle):
    rng = np.random.default_rng(4)
    a = (rng.standard_normal((2, 1500)) + 1j * rng.standard_normal((2, 1500)))
    b = (rng.standard_normal((2, 1500)) + 1j * rng.standard_normal((2, 1500)))
    f = oracle.range_compress
    assert np.allclose(f(a + 2j * b), f(a) + 2j * f(b), atol=1e-9)
    raw = oracle.lfm_echo(2000, 800.0)
    s1 = f(raw)[0]
    s2 = f(np.roll(raw, 100, axis=1))[0]
    assert np.argmax(np.abs(s2)) == np.argmax(np.abs(s1)) + 100
