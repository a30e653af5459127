"""Numpy restatement of required-tile marking (bp_project_amd/culling.py and
the marking kernels), pixel by pixel, for small plans in tests.

``reference_masks`` propagates 32 x 32 tile masks from the root down: every
pixel of a marked parent tile whose stencil origin is defined marks the child
tiles its stencil box touches.  ``ok_only`` counts only pixels whose stencil
lies inside the child (the pixels that read anything: the exact necessary
set); ``widen`` widens each box by that many cells (the device widens by 1).
"""
import numpy as np

from bp_project_amd.culling import TILE, crop_tiles
from bp_project_amd.plan import iter_nodes
from bp_project_amd.plan_audit import STENCILS, source_coordinates


def _tiles_of(nr, nu, tile=TILE):
    return -(-nr // tile), -(-nu // tile)


def _pixels_of(mask, nr, nu, tile=TILE):
    pix = np.repeat(np.repeat(mask, tile, 0), tile, 1)[:nr, :nu]
    return np.flatnonzero(pix.ravel())


def _mark(mask, nr, nu, ir, iu, first, taps, widen, ok_only, tile=TILE):
    lo_r, hi_r = ir + first - widen, ir + first + taps - 1 + widen
    lo_u, hi_u = iu + first - widen, iu + first + taps - 1 + widen
    if ok_only:
        keep = (ir + first >= 0) & (ir + first + taps - 1 <= nr - 1) & \
               (iu + first >= 0) & (iu + first + taps - 1 <= nu - 1)
    else:
        keep = (hi_r >= 0) & (lo_r <= nr - 1) & (hi_u >= 0) & (lo_u <= nu - 1)
    lo_r, hi_r = np.clip(lo_r[keep], 0, nr - 1), np.clip(hi_r[keep], 0, nr - 1)
    lo_u, hi_u = np.clip(lo_u[keep], 0, nu - 1), np.clip(hi_u[keep], 0, nu - 1)
    for a0, a1, b0, b1 in {(int(a), int(b), int(c), int(d))
                           for a, b, c, d in zip(lo_r // tile, hi_r // tile, lo_u // tile,
                                                 hi_u // tile)}:
        mask[a0:a1 + 1, b0:b1 + 1] = True


def grid_root_tiles(plan, grid, interp="kaiser8", widen=1, ok_only=False):
    """Root tiles read by the Cartesian evaluation of the crop onto ``grid``."""
    first, taps = STENCILS[interp]
    root = plan.root
    mask = np.zeros(_tiles_of(root.nr, root.nu), bool)
    yy, xx = np.meshgrid(grid.y, grid.x, indexing="ij")
    dx = xx.ravel() - plan.origin[0]
    dy = yy.ravel() - plan.origin[1]
    s2 = dx * dx + dy * dy
    ok = s2 > 0
    rr = np.sqrt(s2[ok] + plan.origin[2] ** 2)
    uu = dy[ok] / np.sqrt(s2[ok])
    ir = np.floor((rr - plan.r0) / plan.dr).astype(np.int64)
    iu = np.floor((uu - plan.u0) / plan.du).astype(np.int64)
    # stencils in crop coordinates, then shifted into the root
    lo_r, hi_r = ir + first - widen, ir + first + taps - 1 + widen
    lo_u, hi_u = iu + first - widen, iu + first + taps - 1 + widen
    if ok_only:
        keep = (ir + first >= 0) & (ir + first + taps - 1 <= plan.nr - 1) & \
               (iu + first >= 0) & (iu + first + taps - 1 <= plan.nu - 1)
    else:
        keep = (hi_r >= 0) & (lo_r <= plan.nr - 1) & (hi_u >= 0) & (lo_u <= plan.nu - 1)
    lo_r = np.clip(lo_r[keep], 0, plan.nr - 1) + plan.crop_r[0]
    hi_r = np.clip(hi_r[keep], 0, plan.nr - 1) + plan.crop_r[0]
    lo_u = np.clip(lo_u[keep], 0, plan.nu - 1) + plan.crop_u[0]
    hi_u = np.clip(hi_u[keep], 0, plan.nu - 1) + plan.crop_u[0]
    for a0, a1, b0, b1 in {(int(a), int(b), int(c), int(d))
                           for a, b, c, d in zip(lo_r // TILE, hi_r // TILE, lo_u // TILE,
                                                 hi_u // TILE)}:
        mask[a0:a1 + 1, b0:b1 + 1] = True
    return mask


def reference_masks(plan, interp="kaiser8", widen=1, ok_only=False, root_mask=None, tile=TILE):
    """{node index: bool tile mask} from the root crop (or ``root_mask``) down;
    ``tile`` = 1 gives pixel masks (root: the crop pixels)."""
    first, taps = STENCILS[interp]
    if root_mask is None and tile == 1:
        root_mask = np.zeros((plan.root.nr, plan.root.nu), bool)
        root_mask[slice(*plan.crop_r), slice(*plan.crop_u)] = True
    masks = {plan.root.index: crop_tiles(plan) if root_mask is None else root_mask}
    for node in iter_nodes(plan.root):  # pre-order: parents before children
        if node.is_leaf:
            continue
        pixels = _pixels_of(masks[node.index], node.nr, node.nu, tile)
        for child in (node.left, node.right):
            mask = np.zeros(_tiles_of(child.nr, child.nu, tile), bool)
            qr, qu, defined = source_coordinates(node, child, pixels)
            ir = np.floor(qr[defined]).astype(np.int64)
            iu = np.floor(qu[defined]).astype(np.int64)
            _mark(mask, child.nr, child.nu, ir, iu, first, taps, widen, ok_only, tile)
            masks[child.index] = mask
    return masks
