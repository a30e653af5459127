"""Plan-support audit (fast_bp/resident_plan_audit.py) on the CPU.

Pinned by brute force (each required pixel's stencil walked in plain Python
on a small plan), by invariants of the radial guard extension, and by a
planted defect (a node shrunk under its parent's stencils must fail)."""
import numpy as np
import pytest

from bp_project_amd.gpu_backend import build_ffbp_plan
from bp_project_amd.plan import iter_nodes
from bp_project_amd.plan_audit import (STENCILS, audit_support, extend_plan_radial_guard,
                                       minimum_guard, source_coordinates)
from bp_project_amd.scenes import synthetic_kat_scene


def small_plan(leaf=128, oversample=3.):
    c = synthetic_kat_scene()
    return build_ffbp_plan(c["pos"], c["fc"], (700, 760, -.3, -.1), leaf_size=leaf,
                           range_spacing=.5, angular_oversample=oversample)


def brute_force(plan, interp):
    """Per pixel: floor indices, validity and stencil union, in Python loops."""
    first, taps = STENCILS[interp]
    edges = []

    def walk(node, required):
        if node.is_leaf:
            return
        for child in (node.left, node.right):
            need = set()
            invalid = 0
            for p in sorted(required):
                qr, qu, ok = source_coordinates(node, child, [p[0] * node.nu + p[1]])
                if not ok[0]:
                    invalid += 1
                    continue
                i, j = int(np.floor(qr[0])), int(np.floor(qu[0]))
                if i + first < 0 or j + first < 0 or i + first + taps > child.nr or \
                        j + first + taps > child.nu:
                    invalid += 1
                    continue
                for a in range(taps):
                    for b in range(taps):
                        need.add((i + first + a, j + first + b))
            edges.append((node.index, child.index, len(required), invalid, len(need)))
            walk(child, need)

    root = {(r, u) for r in range(*plan.crop_r) for u in range(*plan.crop_u)}
    walk(plan.root, root)
    return edges


@pytest.mark.parametrize("interp", ["lagrange4", "kaiser8"])
def test_audit_matches_brute_force(interp):
    plan = small_plan(leaf=128)
    rep = audit_support(plan, interp)
    got = [(r["destination_index"], r["source_index"], r["required_destination_sites"],
            r["invalid_required_stencils"], r["required_source_sites"]) for r in rep["records"]]
    assert got == brute_force(plan, interp)
    assert rep["checked_edges"] == 2 * plan.counts["merges"]
    assert rep["records"][0]["required_destination_sites"] == plan.nr * plan.nu
    assert rep["passed"] == (rep["invalid_required_stencils"] == 0)


def test_guard_extension_invariants():
    plan = small_plan(leaf=64)
    ext = extend_plan_radial_guard(plan, 5)
    assert ext.crop_r == (plan.crop_r[0] + 5, plan.crop_r[1] + 5)
    for a, b in zip(iter_nodes(plan.root), iter_nodes(ext.root)):
        assert b.nr == a.nr + 10 and b.r0 == pytest.approx(a.r0 - 5 * a.dr, abs=1e-12)
        assert b.nu == a.nu and b.u0 == a.u0
    assert ext.counts["merged_pixels"] == sum(
        2 * n.nr * n.nu for n in iter_nodes(ext.root) if not n.is_leaf)
    ext2 = extend_plan_radial_guard(ext, 3)
    assert ext2.guard_extension["cumulative_extra_cells_per_side"] == 8
    # the crop's slant ranges are unchanged
    r_before = plan.root.r0 + plan.root.dr * np.arange(*plan.crop_r)
    r_after = ext2.root.r0 + ext2.root.dr * np.arange(*ext2.crop_r)
    assert np.allclose(r_before, r_after, atol=1e-9)
    with pytest.raises(ValueError):
        extend_plan_radial_guard(plan, -1)
    with pytest.raises(ValueError):
        extend_plan_radial_guard(plan, True)


def test_planted_defect_fails_and_guards_fix_it():
    """Shrinking one child radially by 12 cells on each side puts parent
    stencils outside it; a wide enough guard extension repairs the plan."""
    plan = small_plan(leaf=64)
    assert audit_support(plan, "lagrange4")["passed"]
    bad = extend_plan_radial_guard(plan, 0)
    child = bad.root.left
    child.r0 += 12 * child.dr
    child.nr -= 24
    rep = audit_support(bad, "lagrange4")
    assert not rep["passed"] and rep["records"][0]["invalid_required_stencils"] > 0
    assert rep["records"][0]["invalid_example"] is not None
    # kaiser8 needs at least as much support as lagrange4 on every edge
    lag = audit_support(plan, "lagrange4")["records"]
    kai = audit_support(plan, "kaiser8")["records"]
    for a, b in zip(lag, kai):
        assert b["invalid_required_stencils"] >= a["invalid_required_stencils"]
        assert b["required_source_sites"] >= a["required_source_sites"]
    g = minimum_guard(plan, "kaiser8", limit=16)
    assert g is not None
    assert audit_support(extend_plan_radial_guard(plan, g), "kaiser8")["passed"]


def test_source_coordinates_and_validity_against_oracle_regrid(oracle):
    """4-point Lagrange reproduces linear functions exactly, so regridding a
    child image holding its own indices (row + j col) with the oracle's
    regrid (k = 0: no phase) returns qr + j qu wherever the stencil is valid
    and 0 elsewhere: an independent pin of the audit's geometry and rule."""
    plan = small_plan(leaf=64)
    parent, child = plan.root, plan.root.left
    ii, jj = np.meshgrid(np.arange(child.nr), np.arange(child.nu), indexing="ij")
    img = (ii + 1j * jj).astype(np.complex128)
    out = oracle.regrid(img, child.origin, child.r0, child.dr, child.u0, child.du, parent.origin,
                        parent.r0, parent.dr, parent.nr, parent.u0, parent.du, parent.nu, 0.0,
                        parent.side)
    qr, qu, defined = source_coordinates(parent, child, np.arange(parent.nr * parent.nu))
    sr, su = np.floor(qr), np.floor(qu)
    valid = defined & (sr >= 1) & (su >= 1) & (sr <= child.nr - 3) & (su <= child.nu - 3)
    got = out.ravel()
    assert valid.any() and (~valid).any()
    assert np.allclose(got[valid].real, qr[valid], atol=1e-7)
    assert np.allclose(got[valid].imag, qu[valid], atol=1e-7)
    assert np.all(got[~valid] == 0)
