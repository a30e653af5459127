"""Pins of the oracle's interpolation primitives by closed forms and brute
force (fast_bp/ffbp.py:17-40; linear and kaiser8 have no reference
counterpart and are checked against their definitions)."""
import numpy as np
import pytest


def test_cubic_weights_closed_form(oracle):
    assert oracle.cubic_weights(0.0) == (0.0, 1.0, 0.0, 0.0)
    for t in np.linspace(0, 1, 17):
        w = np.array(oracle.cubic_weights(t))
        assert abs(w.sum() - 1) < 1e-14
        # exact reproduction of every cubic through taps -1, 0, 1, 2
        for coeffs in ([1, 0, 0, 0], [0.3, -2, 0.5, 1.25], [0, 0, 0, 1]):
            poly = np.polynomial.Polynomial(coeffs)
            taps = poly(np.array([-1.0, 0.0, 1.0, 2.0]))
            assert abs(np.dot(w, taps) - poly(t)) < 1e-12
    np.testing.assert_allclose(oracle.np_cubic_weights(0.37), oracle.cubic_weights(0.37), rtol=0, atol=1e-15)


def test_sample1_boundary_rule_and_cubic(oracle):
    n = 12
    x = np.arange(n, dtype=np.float64)
    row = ((0.5 * x ** 3 - x + 2) + 1j * (x ** 2)).astype(np.complex64)
    # valid indices 1 <= floor(q) <= n - 3 (ffbp.py:27)
    assert oracle.sample1(row, 0.999) == 0
    assert oracle.sample1(row, n - 2.0) == 0
    assert oracle.sample1(row, 1.0) == complex(row[1])
    for q in (1.0, 1.25, 4.5, 8.99):
        exact = (0.5 * q ** 3 - q + 2) + 1j * q ** 2
        assert abs(oracle.sample1(row, q) - exact) < 1e-4 * abs(exact)
        assert abs(oracle.sample1(row, q) - oracle.np_sample1(row, np.array([q]))[0]) < 1e-12


def test_sample2_boundary_rule_and_separable_cubic(oracle):
    nr, nu = 9, 11
    r = np.arange(nr, dtype=np.float64)[:, None]
    u = np.arange(nu, dtype=np.float64)[None, :]
    img = ((r ** 2 - u) + 1j * (r * u ** 3 / 50)).astype(np.complex64)
    assert oracle.sample2(img, 0.5, 4.0) == 0
    assert oracle.sample2(img, 4.0, nu - 2.0) == 0
    for qr, qu in ((1.0, 1.0), (3.3, 4.7), (5.9, 7.99)):
        exact = (qr ** 2 - qu) + 1j * (qr * qu ** 3 / 50)
        got = oracle.sample2(img, qr, qu)
        assert abs(got - exact) < 2e-4 * max(1.0, abs(exact))
        assert abs(got - oracle.np_sample2(img, np.array([qr]), np.array([qu]))[0]) < 1e-11
    # complex128 images are read without rounding
    img128 = img.astype(np.complex128) * (1 + 1e-9)
    assert oracle.sample2(img128, 3.3, 4.7) != oracle.sample2(img, 3.3, 4.7)


def test_linear_interpolator_definition(oracle):
    row = (np.arange(6) * (1 + 2j)).astype(np.complex64)
    assert oracle.sample1(row, 0.0, "linear") == 0
    assert abs(oracle.sample1(row, 2.25, "linear") - 2.25 * (1 + 2j)) < 1e-12
    assert oracle.sample1(row, 5.0, "linear") == 0          # needs i + 1 <= len - 1
    assert oracle.sample1(row, -0.5, "linear") == 0


def test_kaiser_weights_definition(oracle):
    w0 = oracle.kaiser_weights(0.0)
    np.testing.assert_allclose(w0, [0, 0, 0, 1, 0, 0, 0, 0], atol=1e-15)
    for t in (0.1, 0.5, 0.77):
        w = oracle.kaiser_weights(t)
        assert abs(w.sum() - 1) < 1e-14
        # symmetric kernel: weights at t and 1 - t mirror each other
        np.testing.assert_allclose(oracle.kaiser_weights(1 - t), w[::-1], atol=1e-15)
    # band-limited tone (0.2 cycles/sample) reproduced to the window's accuracy
    n = 64
    x = np.arange(n)
    row = np.exp(2j * np.pi * 0.2 * x).astype(np.complex64)
    for q in (20.3, 31.5, 40.9):
        got = oracle.sample1(row, q, "kaiser8")
        assert abs(got - np.exp(2j * np.pi * 0.2 * q)) < 5e-3
    assert oracle.sample1(row, 2.5, "kaiser8") == 0         # needs i >= 3
    assert oracle.sample1(row, n - 4.5, "kaiser8") != 0     # i + 4 = len - 1 is valid
    assert oracle.sample1(row, n - 4.0, "kaiser8") == 0     # needs i + 4 <= len - 1


def test_kaiser_table_matches_oracle_kernel(oracle):
    from bp_project_amd.gpu_backend import kaiser_table
    table = kaiser_table()
    L = oracle.lib()
    for i in (0, 1, 100, 2048, 3000, 6000, 8191, 8192, 8193):
        assert abs(table[i] - L.orc_kaiser_sinc(i / 2048.0)) < 1e-14
        assert abs(table[i] - L.orc_kaiser_table_value(i)) < 1e-14
