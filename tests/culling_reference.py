"""Host restatement of required-tile marking (bp_project_amd/culling.py and
the marking kernels) for small plans in tests.

``reference_masks`` propagates 32 x 32 tile masks from the root down with the
plan audit's support walk (bp_project_amd.plan_audit.walk_support): every
pixel of a marked parent tile whose stencil origin is defined marks the child
tiles its stencil box touches.  ``ok_only`` counts only pixels whose stencil
lies inside the child (the pixels that read anything: the exact necessary
set); ``widen`` widens each box by that many cells (the device widens by 1).
``grid_root_tiles`` restates the root tiles a Cartesian evaluation reads.
"""
import numpy as np

from bp_project_amd.culling import TILE, crop_tiles
from bp_project_amd.plan_audit import STENCILS, crop_mask, walk_support


def _tiles_of(nr, nu, tile=TILE):
    return -(-nr // tile), -(-nu // tile)


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
    """{node index: bool tile mask} from the root crop (or ``root_mask``) down,
    by the plan audit's support walk; ``tile`` = 1 gives pixel masks."""
    if root_mask is None:
        root_mask = crop_mask(plan) if tile == 1 else crop_tiles(plan)
    masks = {plan.root.index: root_mask}
    for e in walk_support(plan, interp, root_mask, tile=tile, widen=widen, inside_only=ok_only):
        masks[e.child.index] = e.child_mask
    return masks
