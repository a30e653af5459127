"""Support of an FFBP plan: which node pixels the output depends on, and
whether every one of them finds its whole interpolation stencil inside the
child image it reads.

``walk_support`` is the one engine.  From a mask of required root pixels it
visits the tree edges (parent, child) depth first -- an edge, then the whole
subtree under that child -- and maps the parent's required pixels into the
child with the regrid geometry (fast_bp/ffbp.py:79-95).  Each pixel's
stencil (4x4 Lagrange, fast_bp/ffbp.py:32-40, or this build's 8x8
Kaiser-sinc) either lies inside the child -- then its cells become required
child pixels -- or it does not, and the reference's sample2 reads zeros there
(ffbp.py:35).  At pixel granularity with only inside stencils propagated the
walk gives the exact necessary sets; at the tile granularity of the culled
schedules, with boxes widened, it gives the conservative tile masks that the
device marking must contain (tests/culling_reference.py).

``audit_support`` (the counterpart of fast_bp/resident_plan_audit.py:60-138;
its report keys are kept for callers that read them) runs the walk from the
root crop and reports, per edge, the required pixels whose stencils fall
outside the child: a plan that passes has guard cells wide enough for that
interpolator.  ``extend_plan_radial_guard`` widens every node radially
(resident_plan_audit.py:15-61), ``minimum_guard`` searches the smallest
widening that passes.

Host numpy code: the audit is a property of the FP64 plan.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass

import numpy as np

from .plan import iter_nodes

# interpolator -> (offset of the first tap from floor(q), taps per axis)
STENCILS = {"lagrange4": (-1, 4), "kaiser8": (-3, 8)}


def source_coordinates(destination, source, flat_index):
    """(qr, qu, defined): fractional row and column in ``source`` that the
    regrid samples for destination pixels ``flat_index`` (row-major), and
    whether the point is off the source origin (a zero ground distance has
    no angle).  Geometry of fast_bp/ffbp.py:79-95 in FP64."""
    row, col = np.divmod(np.asarray(flat_index, np.int64), destination.nu)
    o, so = destination.origin, source.origin
    r = destination.r0 + row * destination.dr
    u = destination.u0 + col * destination.du
    ground = np.sqrt(np.maximum(0.0, r * r - o[2] * o[2]))
    ex = o[0] + destination.side * ground * np.sqrt(np.maximum(0.0, 1.0 - u * u)) - so[0]
    ey = o[1] + ground * u - so[1]
    g = np.sqrt(ex * ex + ey * ey)
    with np.errstate(divide="ignore", invalid="ignore"):
        qr = (np.sqrt(g * g + so[2] * so[2]) - source.r0) / source.dr
        qu = (ey / g - source.u0) / source.du
    return qr, qu, (g != 0.0) & np.isfinite(qr) & np.isfinite(qu)


@dataclass
class EdgeSupport:
    parent: object          # plan node
    child: object
    required: int           # required parent pixels
    outside: int            # of those, stencils not wholly inside the child (or undefined)
    margin: int | None      # least distance (cells) from a defined stencil to the child edge
    first_outside: dict | None
    child_mask: np.ndarray  # bool [ceil(nr / tile), ceil(nu / tile)] of the child


def _pixels(mask, nr, nu, tile):
    if tile == 1:
        return np.flatnonzero(mask.ravel())
    full = np.repeat(np.repeat(mask, tile, 0), tile, 1)[:nr, :nu]
    return np.flatnonzero(full.ravel())


def _edge(parent, child, mask, stencil, tile, widen, inside_only, batch):
    first, taps = stencil
    nr, nu = child.nr, child.nu
    out = np.zeros((-(-nr // tile), -(-nu // tile)), bool)
    pix = _pixels(mask, parent.nr, parent.nu, tile)
    outside, margin, example = 0, None, None
    span = (taps + 2 * widen + tile - 2) // tile + 1      # tiles a box can cross per axis
    for s in range(0, len(pix), batch):
        p = pix[s:s + batch]
        qr, qu, defined = source_coordinates(parent, child, p)
        lo_r = np.floor(np.where(defined, qr, -1.0)).astype(np.int64) + first
        lo_u = np.floor(np.where(defined, qu, -1.0)).astype(np.int64) + first
        room = np.minimum(np.minimum(lo_r, lo_u),
                          np.minimum(nr - taps - lo_r, nu - taps - lo_u))
        inside = defined & (room >= 0)
        outside += int(np.count_nonzero(~inside))
        if room.size:
            m = int(room.min())
            margin = m if margin is None else min(margin, m)
        if example is None and not inside.all():
            k = int(np.flatnonzero(~inside)[0])
            example = {"destination_flat_index": int(p[k]), "qr": float(qr[k]),
                       "qu": float(qu[k])}
        keep = inside if inside_only else defined
        a0, b0 = lo_r[keep] - widen, lo_u[keep] - widen
        a1, b1 = a0 + taps - 1 + 2 * widen, b0 + taps - 1 + 2 * widen
        hit = (a1 >= 0) & (a0 < nr) & (b1 >= 0) & (b0 < nu)
        t0r, t1r = np.clip(a0[hit], 0, nr - 1) // tile, np.clip(a1[hit], 0, nr - 1) // tile
        t0u, t1u = np.clip(b0[hit], 0, nu - 1) // tile, np.clip(b1[hit], 0, nu - 1) // tile
        for i in range(span):
            for j in range(span):
                out[np.minimum(t0r + i, t1r), np.minimum(t0u + j, t1u)] = True
    return EdgeSupport(parent, child, len(pix), outside, margin, example, out)


def walk_support(plan, interp, root_mask, *, tile=1, widen=0, inside_only=True,
                 batch=1 << 16):
    """Yield an ``EdgeSupport`` per tree edge, depth first (an edge, then the
    subtree below its child), starting from ``root_mask`` (bool, per pixel
    when ``tile`` is 1, else per tile x tile block of the root).  A child's
    mask marks the cells (``tile`` 1) or tiles that the stencils of its
    parent's required pixels cover, each stencil box widened by ``widen``
    cells and clipped to the child; ``inside_only`` propagates only stencils
    wholly inside the child (those that read anything)."""
    if interp not in STENCILS:
        raise ValueError(f"interp must be one of {tuple(STENCILS)}")
    stencil = STENCILS[interp]

    def visit(node, mask):
        if node.is_leaf:
            return
        for child in (node.left, node.right):
            e = _edge(node, child, mask, stencil, tile, widen, inside_only, batch)
            yield e
            yield from visit(child, e.child_mask)

    yield from visit(plan.root, np.asarray(root_mask, bool))


def crop_mask(plan):
    """Pixel mask of the root crop (the output region of form_ffbp)."""
    m = np.zeros((plan.root.nr, plan.root.nu), bool)
    m[slice(*plan.crop_r), slice(*plan.crop_u)] = True
    return m


def audit_support(plan, interp="lagrange4", batch_size=65536, sites=None):
    """Per-edge support audit of ``plan`` from the root crop down.  ``sites``
    (a dict) receives each node's required-pixel mask by node index.
    Report keys follow fast_bp/resident_plan_audit.py:60-138."""
    taps = STENCILS.get(interp, (0, 0))[1]
    root = crop_mask(plan)
    if sites is not None:
        sites[plan.root.index] = root
    records = []
    for e in walk_support(plan, interp, root, batch=batch_size):
        if sites is not None:
            sites[e.child.index] = e.child_mask
        records.append({
            "destination_index": int(e.parent.index),
            "source_index": int(e.child.index),
            "required_destination_sites": e.required,
            "invalid_required_stencils": e.outside,
            "minimum_integer_support_margin_cells": e.margin,
            "required_source_sites": int(np.count_nonzero(e.child_mask)),
            "invalid_example": e.first_outside,
        })
    bad = sum(1 for r in records if r["invalid_required_stencils"])
    return {
        "method": f"host FP64 plan geometry; {taps}x{taps} {interp} stencils of every "
                  "required pixel, propagated from the root crop to the leaves",
        "interp": interp,
        "passed": bad == 0,
        "checked_edges": len(records),
        "invalid_required_stencils": int(sum(r["invalid_required_stencils"] for r in records)),
        "edges_with_invalid_required_stencils": bad,
        "records": records,
        "limitations": [
            "merge and evaluation stencils only: the leaves' range-sample support and the "
            "interpolation error are not audited",
            "floor() of FP64 coordinates decides each stencil; the device computes the same "
            "coordinates and can differ only within rounding of an integer",
            "every stencil cell counts as required, including cells whose weight is zero",
            "outside stencils are counted once per edge, not per output pixel",
        ],
    }


def extend_plan_radial_guard(plan, extra_cells):
    """Copy of ``plan`` whose nodes have ``extra_cells`` more range cells on
    each side: every node's r0 moves down by that many steps and nr grows by
    twice as many; angles, the output grid and every pixel's position are
    unchanged, and the root crop moves up by ``extra_cells`` rows."""
    if isinstance(extra_cells, bool) or int(extra_cells) != extra_cells or extra_cells < 0:
        raise ValueError("extra_cells must be a nonnegative integer")
    n = int(extra_cells)
    out = copy.deepcopy(plan)
    leaves = merges = leaf_pp = merged = nodes = 0
    for node in iter_nodes(out.root):
        node.r0 -= n * node.dr
        node.nr += 2 * n
        nodes += 1
        if node.is_leaf:
            leaves += 1
            leaf_pp += node.nr * node.nu * (node.b - node.a)
        else:
            merges += 1
            merged += 2 * node.nr * node.nu
    out.crop_r = (int(plan.crop_r[0] + n), int(plan.crop_r[1] + n))
    for name in ("r0", "dr", "nr", "u0", "du", "nu", "side", "crop_u"):
        if getattr(out, name) != getattr(plan, name):
            raise AssertionError(f"output grid changed: {name}")
    before = getattr(plan, "guard_extension", {}).get("cumulative_extra_cells_per_side", 0)
    out.counts = dict(plan.counts, leaves=leaves, merges=merges, nodes=nodes,
                      leaf_pixel_pulse_products=leaf_pp, merged_pixels=merged,
                      shape=[int(out.nr), int(out.nu)],
                      direct_pixel_pulse_products=int(out.nr * out.nu
                                                      * (out.root.b - out.root.a)))
    out.guard_extension = {
        "extra_cells_per_side": n,
        "cumulative_extra_cells_per_side": int(before + n),
        "changed_axes": "range only, both ends of every node",
        "output_grid_unchanged": True,
        "root_crop_r_before": list(plan.crop_r),
        "root_crop_r_after": list(out.crop_r),
        "root_shape_before": [int(plan.root.nr), int(plan.root.nu)],
        "root_shape_after": [int(out.root.nr), int(out.root.nu)],
    }
    return out


def minimum_guard(plan, interp="kaiser8", limit=32, step=2):
    """Smallest radial widening (cells per side, a multiple of ``step``, at
    most ``limit``) whose plan passes ``audit_support``, or None."""
    for extra in range(0, limit + 1, step):
        if audit_support(extend_plan_radial_guard(plan, extra) if extra else plan,
                         interp)["passed"]:
            return extra
    return None
