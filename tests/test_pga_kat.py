"""Phase-gradient autofocus caller (fast_bp/autofocus.py:20-79) on the CPU.

The host half (``phase_gradient_update``) is fed the oracle's FP64
per-(target, pulse) contributions, rounded to complex64 as the reference's
Numba ``target_contributions`` stores them, on the deterministic scene of
fast_bp/validate_autofocus.py; its statistics must reproduce the numbers the
reference published in results/autofocus_validation.json.
"""
import numpy as np
import pytest

from bp_project_amd.autofocus import phase_gradient_update, pick_targets
from bp_project_amd.scenes import autofocus_kat_scene


@pytest.fixture(scope="module")
def kat(oracle):
    s = autofocus_kat_scene()
    xi = oracle.target_contributions(s["data"], s["pos"], s["targets"], s["r_start"],
                                     s["data_dr"], s["k"]).astype(np.complex64)
    return s, xi


def rmse_xy(e):
    return float(np.sqrt(np.mean(e[:, :2] ** 2)))


def test_phase_gradient_update_reproduces_reference_kat(kat, golden):
    s, xi = kat
    ref = golden["autofocus_validation"]
    delta, stats = phase_gradient_update(xi, s["pos"], s["targets"], s["k"], s["prf"], s["cutoff"])
    assert stats["targets"] == ref["targets"] == 4
    assert stats["cutoff_hz"] == ref["cutoff_hz"]
    np.testing.assert_allclose(stats["median_gradient_coherence"],
                               ref["median_gradient_coherence"], rtol=1e-9)
    np.testing.assert_allclose(stats["delta_rms_m"], ref["delta_rms_m"], rtol=1e-8, atol=1e-15)
    np.testing.assert_allclose(stats["delta_max_m"], ref["delta_max_m"], rtol=1e-8, atol=1e-15)
    before = rmse_xy(s["err"])
    after = rmse_xy(s["err"] + delta)
    np.testing.assert_allclose(before, ref["initial_coordinate_rmse_m"], rtol=1e-12)
    np.testing.assert_allclose(after, ref["corrected_coordinate_rmse_m"], rtol=1e-8)
    assert after < .35 * before  # the reference's own acceptance (validate_autofocus.py:21)
    assert delta[0, 2] == 0 and not delta[:, 2].any()


def test_phase_gradient_update_sign_and_null(kat):
    """A trajectory without error yields (almost) no update; flipping the
    error flips the update."""
    s, _ = kat
    import oracle as O
    args = (s["r_start"], s["data_dr"], s["k"])
    xi0 = O.target_contributions(s["data"], s["truth"], s["targets"], *args).astype(np.complex64)
    d0, _ = phase_gradient_update(xi0, s["truth"], s["targets"], s["k"], s["prf"], s["cutoff"])
    assert rmse_xy(d0) < 0.02 * rmse_xy(s["err"])
    neg = s["truth"] - s["err"]
    xin = O.target_contributions(s["data"], neg, s["targets"], *args).astype(np.complex64)
    dn, _ = phase_gradient_update(xin, neg, s["targets"], s["k"], s["prf"], s["cutoff"])
    assert rmse_xy(-s["err"] + dn) < .35 * rmse_xy(s["err"])


def test_pick_targets_finds_the_point_targets(oracle):
    """Brightest blocks of an FFBP image of the validate.py scene land on its
    three point targets (within a resolution cell); dim blocks are dropped."""
    from bp_project_amd.scenes import synthetic_kat_scene
    c = synthetic_kat_scene()
    res, _ = oracle.form_ffbp(c["data"], c["pos"], c["r_start"], c["data_dr"], c["fc"],
                              c["extent"], leaf_size=128, range_spacing=c["range_spacing"],
                              angular_oversample=c["angular_oversample"], weight=c["weight"])
    points, powers = pick_targets(res, nr_tiles=6, nu_tiles=6)
    assert len(points) == len(powers) and np.all(powers > powers.max() * 10 ** -3.5)
    assert len(points) < 36
    for xyz in c["targets"]:
        d = np.linalg.norm(points[:, :2] - xyz[:2], axis=1).min()
        assert d < 1.0, (xyz, d)
    assert np.all(points[:, 2] == 0)
