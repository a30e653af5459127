"""Required-tile culling of a split-mode FFBP schedule.

The reference sizes every node to the whole root sector as seen from the
node's origin (fast_bp/gpu_backend.py:121-224) and computes every pixel.
Only the pixels that the root's output region depends on matter: a parent
pixel reads a child only through its stencil.  ``required_masks`` finds, for
every node, the 32 x 32 tiles (the tiled merge's tiles) that the output
region can reach -- on the device, with the same geometry the merge and
evaluation kernels use, from the root down (bpmi_ffbp_required_tiles,
bpmi_eval_required_tiles).  A schedule that launches only those tiles
produces the same output region, bit for bit; the other pixels of non-root
nodes are never written.

Output regions:

* ``"crop"`` -- the root crop, i.e. the whole polar image the program
  returns (the reference's output);
* a Cartesian grid (an object with x0, dx, nx, y0, dy, ny) -- the root pixels
  that the split-mode grid evaluation of the crop reads; the crop pixels
  outside that footprint are then not computed.

``tile_selection`` turns the masks into the merge tiles of each parent and
the split tiles (2^tc columns x 1024 / 2^tc rows) of each leaf that touch a
set tile.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .device import DeviceArray, stream_handle, synchronize
from .plan import NI_NR, NI_NU

_vp = _lib.c_void_p
TILE = 32           # rows and columns of a merge tile
SPLIT_TILE = 1024   # pixels of a leaf split tile


def mask_layout(node_int):
    """Per-node tile grids and byte offsets of the mask array:
    (offsets int64 [nodes], tiles_r, tiles_u, total bytes)."""
    node_int = np.asarray(node_int)
    tiles_r = -(-node_int[:, NI_NR].astype(np.int64) // TILE)
    tiles_u = -(-node_int[:, NI_NU].astype(np.int64) // TILE)
    sizes = tiles_r * tiles_u
    offsets = np.zeros(len(sizes), np.int64)
    offsets[1:] = np.cumsum(sizes)[:-1]
    return offsets, tiles_r, tiles_u, int(sizes.sum())


def crop_tiles(plan):
    """Root tile mask (bool [tiles_r, tiles_u]) of the tiles overlapping the crop."""
    root = plan.root
    mask = np.zeros((-(-root.nr // TILE), -(-root.nu // TILE)), bool)
    (r0, r1), (u0, u1) = plan.crop_r, plan.crop_u
    mask[r0 // TILE:(r1 - 1) // TILE + 1, u0 // TILE:(u1 - 1) // TILE + 1] = True
    return mask


def required_masks(flat, plan, node_table, geo_table, region="crop", merge_interp="kaiser8",
                   eval_interp="kaiser8", stream=None):
    """Boolean tile masks [tiles_r, tiles_u] of every node (list by node
    index).  ``node_table``/``geo_table``: the flat plan's node_int/node_geo
    on the device.  Synchronises the device."""
    offsets, tiles_r, tiles_u, total = mask_layout(flat.node_int)
    masks = DeviceArray.zeros((total,), np.uint8, stream=stream)
    root = plan.root.index
    hs = stream_handle(stream)
    if isinstance(region, str):
        if region != "crop":
            raise ValueError("region must be 'crop' or a Cartesian grid")
        host = np.zeros(total, np.uint8)
        host[offsets[root]:offsets[root] + tiles_r[root] * tiles_u[root]] = \
            crop_tiles(plan).ravel()
        masks.set(host, stream=stream)
    else:
        from .gpu_backend import geo_array
        geo = geo_array(plan)
        _lib.call("bpmi_eval_required_tiles", geo.ctypes.data_as(_vp), plan.nr, plan.nu,
                  plan.crop_r[0], plan.crop_u[0], plan.root.nr, plan.root.nu,
                  float(region.x0), float(region.dx), int(region.nx), float(region.y0),
                  float(region.dy), int(region.ny), _lib.interp_code(eval_interp),
                  _vp(masks.ptr + int(offsets[root])), hs)
    d_offsets = DeviceArray.from_host(offsets, np.int64, stream=stream)
    tables = []
    for height in range(len(flat.levels) - 1, 0, -1):
        entries = np.array([(n, t) for n in flat.levels[height]
                            for t in range(int(tiles_r[n] * tiles_u[n]))], np.int32)
        table = DeviceArray.from_host(entries, np.int32, stream=stream)
        tables.append(table)
        _lib.call("bpmi_ffbp_required_tiles", _vp(node_table.ptr), _vp(geo_table.ptr),
                  _vp(d_offsets.ptr), _vp(masks.ptr), _vp(table.ptr), len(entries),
                  _lib.interp_code(merge_interp), hs)
    synchronize()
    host = masks.get().astype(bool)
    return [host[offsets[i]:offsets[i] + tiles_r[i] * tiles_u[i]].reshape(tiles_r[i], tiles_u[i])
            for i in range(len(offsets))]


def merge_tiles_of(mask):
    """Merge tile indices (row-major over the tile grid) whose mask is set."""
    return np.flatnonzero(np.asarray(mask).ravel())


def split_tiles_of(mask, nr, nu, tile_cols_log2):
    """Leaf split-tile indices (tiles of 2^tc columns x 1024 / 2^tc rows, row-major,
    as bpmi_split_tiles_ex counts them) that overlap a set 32 x 32 tile."""
    tc = 1 << int(tile_cols_log2)
    tr = SPLIT_TILE // tc
    rows, cols = -(-int(nr) // tr), -(-int(nu) // tc)
    pixel = np.repeat(np.repeat(np.asarray(mask, bool), TILE, 0), TILE, 1)[:nr, :nu]
    padded = np.zeros((rows * tr, cols * tc), bool)
    padded[:nr, :nu] = pixel
    hit = padded.reshape(rows, tr, cols, tc).any(axis=(1, 3))
    return np.flatnonzero(hit.ravel())


def tile_selection(masks, flat, tile_cols_log2):
    """{node: tile indices} for leaves (split tiles) and parents (merge tiles),
    plus the pixel counts computed per node."""
    tiles, pixels = {}, {}
    for h, nodes in enumerate(flat.levels):
        for n in nodes:
            nr, nu = int(flat.node_int[n, NI_NR]), int(flat.node_int[n, NI_NU])
            if h == 0:
                sel = split_tiles_of(masks[n], nr, nu, tile_cols_log2)
                tc = 1 << int(tile_cols_log2)
                tr = SPLIT_TILE // tc
                cols = -(-nu // tc)
                r0, c0 = (sel // cols) * tr, (sel % cols) * tc
            else:
                sel = merge_tiles_of(masks[n])
                tu = -(-nu // TILE)
                r0, c0 = (sel // tu) * TILE, (sel % tu) * TILE
                tr = tc = TILE
            tiles[n] = sel
            pixels[n] = int(np.sum((np.minimum(r0 + tr, nr) - r0) * (np.minimum(c0 + tc, nu) - c0)))
    return tiles, pixels
