"""CPU structural audit of FFBP regrid support (fast_bp/resident_plan_audit.py),
for the 4-point Lagrange stencil of the reference and the 8-tap Kaiser-sinc
stencil of this implementation.

``audit_support`` propagates, edge by edge from the root crop down to the
leaves, every source sample a required destination pixel's stencil may read
(FP64 polar geometry of fast_bp/ffbp.py:79-95), and counts the required
stencils that leave their source image -- those read zeros (the validity rule
of ``sample2``, ffbp.py:35), so a failed audit means the plan's guard cells
are too narrow for that interpolator.  Zero-weight stencil samples are
included (conservative), as in the reference.
``extend_plan_radial_guard`` pads every node radially (resident_plan_audit.py:
14-57), keeping all sample coordinates and the output grid.

No device code: the audit is a property of the FP64 plan.
"""
from __future__ import annotations

import copy

import numpy as np

from .plan import iter_nodes

# stencil offsets and validity margin of each interpolator: taps at
# floor(q) + FIRST .. floor(q) + FIRST + TAPS - 1, valid when every tap is inside
STENCILS = {"lagrange4": (-1, 4), "kaiser8": (-3, 8)}


def extend_plan_radial_guard(plan, extra_cells):
    """Deep copy of ``plan`` with ``extra_cells`` more radial cells on both
    sides of every node; node sample coordinates and the output grid are kept
    and the root crop moves by ``extra_cells``."""
    if isinstance(extra_cells, bool) or int(extra_cells) != extra_cells or extra_cells < 0:
        raise ValueError("extra_cells must be a nonnegative integer")
    extra = int(extra_cells)
    result = copy.deepcopy(plan)
    counts = {"leaves": 0, "merges": 0, "leaf_pixel_pulse_products": 0, "merged_pixels": 0,
              "nodes": 0}
    for node in iter_nodes(result.root):
        node.r0 -= extra * node.dr
        node.nr += 2 * extra
        counts["nodes"] += 1
        if node.is_leaf:
            counts["leaves"] += 1
            counts["leaf_pixel_pulse_products"] += node.nr * node.nu * (node.b - node.a)
        else:
            counts["merges"] += 1
            counts["merged_pixels"] += 2 * node.nr * node.nu
    result.crop_r = tuple(int(v + extra) for v in plan.crop_r)
    counts["shape"] = [int(result.nr), int(result.nu)]
    counts["direct_pixel_pulse_products"] = int(result.nr * result.nu *
                                                (result.root.b - result.root.a))
    result.counts = dict(plan.counts) | counts
    previous = getattr(plan, "guard_extension", {}).get("cumulative_extra_cells_per_side", 0)
    result.guard_extension = {
        "extra_cells_per_side": extra,
        "cumulative_extra_cells_per_side": int(previous + extra),
        "changed_axes": "r only; every node padded on both sides",
        "output_grid_unchanged": True,
        "root_crop_r_before": list(plan.crop_r),
        "root_crop_r_after": list(result.crop_r),
        "root_shape_before": [int(plan.root.nr), int(plan.root.nu)],
        "root_shape_after": [int(result.root.nr), int(result.root.nu)],
    }
    for name in ("r0", "dr", "nr", "u0", "du", "nu", "side", "crop_u"):
        if getattr(result, name) != getattr(plan, name):
            raise AssertionError(name)
    return result


def source_coordinates(destination, source, flat_index):
    """Fractional source-grid coordinates (qr, qu) of destination pixels and
    whether they are defined (off the source origin): ffbp.py:79-95."""
    ir, iu = np.divmod(np.asarray(flat_index, np.int64), destination.nu)
    r = destination.r0 + ir * destination.dr
    u = destination.u0 + iu * destination.du
    gr = np.sqrt(np.maximum(0., r * r - destination.origin[2] ** 2))
    x = destination.origin[0] + destination.side * gr * np.sqrt(np.maximum(0., 1. - u * u))
    y = destination.origin[1] + gr * u
    dx = x - source.origin[0]
    dy = y - source.origin[1]
    g = np.sqrt(dx * dx + dy * dy)
    rr = np.sqrt(g * g + source.origin[2] ** 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        qr = (rr - source.r0) / source.dr
        qu = (dy / g - source.u0) / source.du
    defined = np.isfinite(qr) & np.isfinite(qu) & (g != 0.)
    return qr, qu, defined


def audit_support(plan, interp="lagrange4", batch_size=65536, sites=None):
    """Propagate required stencil support from the root crop to the leaves.
    Returns the reference's report (resident_plan_audit.py:60-138) with the
    interpolator named.  ``sites``: optional dict that receives every node's
    required-site mask (bool [nr, nu]) by node index."""
    if interp not in STENCILS:
        raise ValueError(f"interp must be one of {tuple(STENCILS)}")
    first, taps = STENCILS[interp]
    last = first + taps - 1
    records = []
    root_required = np.zeros((plan.root.nr, plan.root.nu), dtype=bool)
    root_required[slice(*plan.crop_r), slice(*plan.crop_u)] = True

    def propagate(destination, source, required):
        indexes = np.flatnonzero(required.ravel())
        child_required = np.zeros(source.nr * source.nu, dtype=bool)
        invalid_count = 0
        minimum_margin = np.inf
        invalid_example = None
        for begin in range(0, len(indexes), batch_size):
            p = indexes[begin:begin + batch_size]
            qr, qu, defined = source_coordinates(destination, source, p)
            sr = np.floor(np.where(defined, qr, -1.)).astype(np.int64)
            su = np.floor(np.where(defined, qu, -1.)).astype(np.int64)
            margin = np.minimum.reduce((sr + first, su + first, source.nr - 1 - last - sr,
                                        source.nu - 1 - last - su))
            valid = defined & (margin >= 0)
            invalid_count += int(np.count_nonzero(~valid))
            if len(margin):
                minimum_margin = min(minimum_margin, int(margin.min()))
            if invalid_example is None and np.any(~valid):
                k = int(np.flatnonzero(~valid)[0])
                invalid_example = {"destination_flat_index": int(p[k]), "qr": float(qr[k]),
                                   "qu": float(qu[k])}
            sr, su = sr[valid], su[valid]
            for a in range(first, last + 1):
                for b in range(first, last + 1):
                    child_required[(sr + a) * source.nu + su + b] = True
        records.append({
            "destination_index": int(destination.index),
            "source_index": int(source.index),
            "required_destination_sites": int(len(indexes)),
            "invalid_required_stencils": invalid_count,
            "minimum_integer_support_margin_cells": (None if not np.isfinite(minimum_margin)
                                                     else int(minimum_margin)),
            "required_source_sites": int(np.count_nonzero(child_required)),
            "invalid_example": invalid_example,
        })
        return child_required.reshape(source.nr, source.nu)

    def walk(node, required):
        # edges in the reference's order: an edge, then the whole subtree below it
        if sites is not None:
            sites[node.index] = required
        if node.is_leaf:
            return
        for child in (node.left, node.right):
            walk(child, propagate(node, child, required))

    walk(plan.root, root_required)
    failures = [row for row in records if row["invalid_required_stencils"]]
    return {
        "method": f"CPU FP64 geometry; conservative {taps}x{taps} ({interp}) support "
                  "propagation from final crop",
        "interp": interp,
        "passed": not failures,
        "checked_edges": len(records),
        "invalid_required_stencils": int(sum(row["invalid_required_stencils"] for row in records)),
        "edges_with_invalid_required_stencils": len(failures),
        "records": records,
        "limitations": [
            "Audits regrid support only; not leaf raw range support or interpolation accuracy",
            "FP64 index decisions; the GPU's FP64 geometry rounds the same way except within "
            "an ulp of an integer boundary",
            "Conservative inclusion of zero-weight stencil samples can overstate downstream "
            "requirements",
            "Invalid stencil counts are per tree edge, not unique final output pixels",
        ],
    }


def minimum_guard(plan, interp="kaiser8", limit=32, step=2):
    """Smallest extra radial guard (cells per side, multiple of ``step``) whose
    extended plan passes the audit, or None up to ``limit``."""
    for extra in range(0, limit + 1, step):
        candidate = extend_plan_radial_guard(plan, extra) if extra else plan
        if audit_support(candidate, interp)["passed"]:
            return extra
    return None
