"""The FP64 oracle adjoint and the entropy gradient are pinned by central
finite differences (h = 1e-4) of the oracle's forward BP and entropy."""
import numpy as np
import pytest

from bp_project_amd.scenes import stripmap_scene


@pytest.fixture(scope="module")
def small_scene():
    return stripmap_scene(pulses=96, bins=1024, n_targets=4, grid_n=24, grid_spacing=0.5, seed=11,
                          phase_walk_rms=0.3)


def test_entropy_gradient_closed_form(oracle):
    rng = np.random.default_rng(5)
    I = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    E, g = oracle.entropy_and_grad(I)
    h = 1e-6
    for j in (0, 7, 33):
        for direction in (1.0, 1j):
            Ip = I.copy(); Ip[j] += h * direction
            Im = I.copy(); Im[j] -= h * direction
            fd = (oracle.entropy_and_grad(Ip)[0] - oracle.entropy_and_grad(Im)[0]) / (2 * h)
            # dE = Re(conj(g) dI)
            assert abs(fd - np.real(np.conj(g[j]) * direction)) < 1e-7


@pytest.mark.parametrize("interp", ["lagrange4", "linear"])
def test_adjoint_matches_central_differences(oracle, small_scene, interp):
    s = small_scene
    pts = s.grid.points()

    def entropy(phase, pos):
        img = oracle.backproject_points(s.data, pos, pts, s.r_start, s.data_dr, s.k, s.weight,
                                        phase, interp=interp)
        return oracle.entropy_and_grad(img)

    E, g = entropy(s.phase, s.pos)
    gphi, gpos = oracle.backproject_points_adjoint(s.data, s.pos, pts, s.r_start, s.data_dr, s.k,
                                                   s.weight, s.phase, g, interp=interp)
    h = 1e-4
    rng = np.random.default_rng(2)
    for n in rng.choice(s.pulses, 6, replace=False):
        ph_p, ph_m = s.phase.copy(), s.phase.copy()
        ph_p[n] += h
        ph_m[n] -= h
        fd = (entropy(ph_p, s.pos)[0] - entropy(ph_m, s.pos)[0]) / (2 * h)
        assert abs(fd - gphi[n]) <= 1e-4 * np.abs(gphi).max() + 1e-9, (n, fd, gphi[n])
        for axis in range(3):
            p_p, p_m = s.pos.copy(), s.pos.copy()
            p_p[n, axis] += h * 1e-2
            p_m[n, axis] -= h * 1e-2
            fd = (entropy(s.phase, p_p)[0] - entropy(s.phase, p_m)[0]) / (2 * h * 1e-2)
            scale = np.abs(gpos[:, axis]).max()
            assert abs(fd - gpos[n, axis]) <= 2e-3 * scale + 1e-9, (n, axis, fd, gpos[n, axis])
