"""oracle.ffbp_at_points (FFBP evaluated at points, computing only the node
pixels the points depend on) equals evaluate_image(form_ffbp(...)) at the
same points bit for bit, for both storage policies and both merge
interpolators; and a support that is cut short (a planted mistake: stencils
not widened and the Kaiser stencil taken as the 4-tap one) changes values,
so the equality is not vacuous."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def c1_small():
    from bp_project_amd.scenes import workload_scene, polar_extent_for_grid
    s = workload_scene("c1_direct_small", pulses=512, bins=1024, grid_n=128)
    ext = polar_extent_for_grid(s.grid, s.pos.mean(axis=0), margin=2.0)
    return s, ext


KW = dict(leaf_size=16, range_spacing=0.25, angular_oversample=2.5)


def _full(oracle, s, ext, policy, interp, pts):
    res, _ = oracle.form_ffbp(s.data, s.pos, s.r_start, s.data_dr, s.fc, ext, weight=s.weight,
                              phase=s.phase, policy=policy, merge_interp=interp, **KW)
    return oracle.evaluate_image(res.image, res.origin, res.r0, res.dr, res.u0, res.du, pts,
                                 s.k, interp=interp)


@pytest.mark.parametrize("policy", ["fp64", "parity"])
@pytest.mark.parametrize("interp", ["kaiser8", "lagrange4"])
def test_support_restricted_ffbp_equals_full(oracle, c1_small, policy, interp):
    s, ext = c1_small
    pts = s.grid.points()
    sel = np.random.default_rng(3).choice(len(pts), 150, replace=False)
    full = _full(oracle, s, ext, policy, interp, pts[sel])
    vals, info = oracle.ffbp_at_points(s.data, s.pos, s.r_start, s.data_dr, s.fc, ext, pts[sel],
                                       weight=s.weight, phase=s.phase, policy=policy,
                                       merge_interp=interp, **KW)
    assert np.abs(full).max() > 0
    assert np.array_equal(vals, full)
    assert info["nodes"] == 63 and 0 < info["computed_fraction"] < 1


def test_cut_short_support_is_detected(oracle, c1_small, monkeypatch):
    s, ext = c1_small
    pts = s.grid.points()
    sel = np.random.default_rng(4).choice(len(pts), 60, replace=False)
    full = _full(oracle, s, ext, "fp64", "kaiser8", pts[sel])
    monkeypatch.setitem(oracle.STENCIL, "kaiser8", (-1, 4))
    vals, _ = oracle.ffbp_at_points(s.data, s.pos, s.r_start, s.data_dr, s.fc, ext, pts[sel],
                                    weight=s.weight, phase=s.phase, policy="fp64",
                                    merge_interp="kaiser8", block=1, **KW)
    assert not np.array_equal(vals, full)
