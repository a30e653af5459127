"""A merge destination pixel that lies exactly on a child's origin (ground
distance 0) has no angle in that child: the oracle's regrid returns 0 there
(the reference GPU kernels' rule, fast_bp/gpu/kernels.cu:359), and the HIP
merge rejects it through the NaN that rsqrt64(0) produces.  The geometry puts
parent pixel (3, 2) exactly on both children's origin: parent origin
(0, 0, 0), r = 9.25 + 3 * 0.25 = 10, u = 0.5 + 2 * 0.25 = 1, so x = 0 and
y = 10, the children's origin (0, 10, 0).  Every other pixel matches the
oracle; nothing is NaN."""
import numpy as np
import pytest

from bp_project_amd import _lib
from bp_project_amd.device import DeviceArray, synchronize

pytestmark = pytest.mark.gpu
_vp = _lib.c_void_p


@pytest.mark.parametrize("mode", ["split", "fp64", "mixed"])
@pytest.mark.parametrize("interp", ["lagrange4", "kaiser8"])
def test_merge_destination_on_child_origin(oracle, mode, interp):
    from bp_project_amd.gpu_backend import ensure_kaiser_table
    if interp == "kaiser8":
        ensure_kaiser_table()
    rng = np.random.default_rng(11)
    cnr, cnu = 96, 64
    child_geo = np.array([0.0, 10.0, 0.0, 0.5, 0.25, -0.9, 0.03], np.float64)
    parent_geo = np.array([0.0, 0.0, 0.0, 9.25, 0.25, 0.5, 0.25], np.float64)
    nr, nu = 8, 3
    left = (rng.standard_normal((cnr, cnu)) + 1j * rng.standard_normal((cnr, cnu))).astype(
        np.complex64)
    right = (rng.standard_normal((cnr, cnu)) + 1j * rng.standard_normal((cnr, cnu))).astype(
        np.complex64)
    k = 4 * np.pi * 9.6e9 / 299792458.0
    dl, dr_, out = DeviceArray.from_host(left), DeviceArray.from_host(right), DeviceArray.empty(
        (nr, nu))
    _lib.call("bpmi_merge2", _vp(dl.ptr), child_geo.ctypes.data_as(_vp), cnr, cnu, _vp(dr_.ptr),
              child_geo.ctypes.data_as(_vp), cnr, cnu, parent_geo.ctypes.data_as(_vp), nr, nu, 1,
              k, _lib.mode_code(mode), _lib.interp_code(interp), _vp(out.ptr), None)
    synchronize()
    got = out.get()
    ref = sum(oracle.regrid(img, child_geo[:3], *child_geo[3:], parent_geo[:3], parent_geo[3],
                            parent_geo[4], nr, parent_geo[5], parent_geo[6], nu, k, 1,
                            interp=interp) for img in (left, right))
    assert np.all(np.isfinite(got))
    assert got[3, 2] == 0 and ref[3, 2] == 0
    valid = ref != 0
    assert valid.sum() >= 6
    err = np.abs(got - ref)[valid].max() / np.abs(ref[valid]).max()
    print(f"{mode} {interp}: max rel error off the origin {err:.2e}")
    assert err <= 1e-5
    assert np.array_equal(got == 0, ref == 0)
