"""GPU image products (fast_bp/run.py:90-128) against the FP64 oracle on the
validate.py scene: the Cartesian ground image equals evaluate_image at the
same points, and the sampled direct-BP validation matches the oracle's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def polar(oracle):
    from bp_project_amd.ffbp import form_ffbp
    from bp_project_amd.scenes import synthetic_kat_scene
    c = synthetic_kat_scene()
    res, counts = form_ffbp(c["data"], c["pos"], c["r_start"], c["data_dr"], c["fc"], c["extent"],
                            leaf_size=64, range_spacing=c["range_spacing"],
                            angular_oversample=c["angular_oversample"], weight=c["weight"],
                            verbose=False)
    return c, res, counts


@pytest.mark.parametrize("side", [1, -1])
def test_ground_image_vs_oracle_evaluate(oracle, polar, side):
    from bp_project_amd.products import ground_image
    c, res, _ = polar
    res.side = side
    try:
        x = np.arange(700., 1000.01, 0.5)
        y = np.arange(-200., 50.01, 0.5)
        g = ground_image(res, x, y, c["k"])
        assert g.shape == (len(y), len(x)) and g.dtype == np.complex64
        yy, xx = np.meshgrid(y, side * x, indexing="ij")
        pts = np.column_stack([xx.ravel(), yy.ravel(), np.zeros(xx.size)])
        ref = oracle.evaluate_image(res.image, res.origin, res.r0, res.dr, res.u0, res.du, pts,
                                    c["k"]).reshape(g.shape)
        err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
        print(f"side {side}: ground image vs oracle evaluate_image {err:.2e}")
        assert err <= 1e-6
    finally:
        res.side = 1


def test_sampled_direct_validation(oracle, polar, tmp_path):
    from bp_project_amd.products import sample_pixels, sampled_direct_validation, save_products
    c, res, counts = polar
    val, samples = sampled_direct_validation(res, c["data"], c["pos"], c["r_start"], c["data_dr"],
                                             c["k"], c["weight"], c["phase"])
    ir, iu = sample_pixels(res)
    assert val["points"] == len(ir) == 4000
    direct_ref = oracle.backproject_points(c["data"], c["pos"], samples["xyz"], c["r_start"],
                                           c["data_dr"], c["k"], c["weight"], c["phase"])
    e = np.linalg.norm(samples["direct"] - direct_ref) / np.linalg.norm(direct_ref)
    assert e <= 1e-6, e
    fast_ref = res.image[ir, iu] * np.exp(1j * c["k"] * res.r[ir])
    ref_err = np.linalg.norm(fast_ref - direct_ref) / np.linalg.norm(direct_ref)
    np.testing.assert_allclose(val["relative_complex_error"], ref_err, rtol=1e-5)
    print(f"sampled FFBP vs direct: error {val['relative_complex_error']:.4f}, "
          f"coherence {val['complex_coherence']:.5f}")
    assert val["passed"]
    paths = save_products(tmp_path, "kat", res, c["k"], counts=counts, validation=val,
                          samples=samples)
    assert len(paths) == 4
