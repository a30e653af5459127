"""Level-batched, preallocated FFBP (fast_bp/batched_resident_backend.py:23-157).

All leaves run in one launch; one launch per subtree height merges every
parent whose children are ready (both children regridded, each rounded to
complex64, then an FP32 add); a final crop produces the requested extent.
Node images live in one arena addressed by per-node offsets, and the node
geometry/offset/block tables are uploaded once at construction.

In ``split`` mode the leaf launch uses the LDS-tiled kernel with each leaf's
origin as the geometry reference (1024-pixel tiles of 2^tc columns, tc from
``leaf_tile_width`` unless given; block table of (node, tile) pairs); the other modes use one thread per pixel with blocks of
256 pixels, like the reference's 128-pixel block table.

``cull="crop"`` (or a Cartesian grid) launches only the tiles that the root
crop (or the grid evaluation's footprint in it) depends on, found on the
device by ``culling.required_masks``; the output region is bitwise
unchanged and the pixels of other tiles stay zero.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray, synchronize
from .gpu_backend import leaf_tile_width
from .plan import NI_NR, NI_NU, flatten_plan
from .resident_backend import ResidentProgram

_vp = _lib.c_void_p

MERGE_BLOCK = 256


def split_tiles(nr, nu, tile_cols_log2=6):
    t = ctypes.c_int()
    _lib.call("bpmi_split_tiles_ex", int(nr), int(nu), int(tile_cols_log2), ctypes.byref(t))
    return t.value


def merge_tiles(nr, nu):
    """Number of 32 x 32 parent tiles of the split-mode tiled merge."""
    t = ctypes.c_int()
    _lib.call("bpmi_merge_tiles", int(nr), int(nu), ctypes.byref(t))
    return t.value


XCDS = 8  # MI355X: workgroup i of a launch runs on accelerator die i % 8


def xcd_order(entries, xcds=XCDS, pad=(-1, 0)):
    """Reorder a block table so that the hardware's round-robin placement of
    workgroups on dies (workgroup i -> die i % xcds) gives each die one
    contiguous run of the original table: die x executes entries
    [x * n / xcds, (x + 1) * n / xcds) in order, so neighbouring tiles, whose
    child windows overlap, share that die's L2.  Padding entries (node -1)
    fill the last row; the kernel skips them."""
    n = len(entries)
    per = -(-n // xcds)
    out = []
    for i in range(per):
        for x in range(xcds):
            j = x * per + i
            out.append(entries[j] if j < n else pad)
    return out


def merge_block_table(node_int, nodes, mode, tiles=None):
    """(node, tile) pairs for split mode (tiled merges, in die order; only
    ``tiles[node]`` when given), else (node, first pixel) pairs of
    MERGE_BLOCK pixels."""
    if mode == "split":
        return xcd_order([(int(n), int(t)) for n in nodes
                          for t in (tiles[n] if tiles is not None else
                                    range(merge_tiles(node_int[n, NI_NR], node_int[n, NI_NU])))])
    return [(int(n), s) for n in nodes
            for s in range(0, int(node_int[n, NI_NR]) * int(node_int[n, NI_NU]), MERGE_BLOCK)]


def enqueue_merges(node_int_ptr, node_geo_ptr, offsets_ptr, arena_ptr, blocks_ptr, nblocks, k,
                   mode, merge_interp, stream):
    """One merge launch over a block table built by ``merge_block_table``."""
    if mode == "split":
        _lib.call("bpmi_ffbp_merges_tiled", _vp(node_int_ptr), _vp(node_geo_ptr),
                  _vp(offsets_ptr), _vp(arena_ptr), _vp(blocks_ptr), int(nblocks), float(k),
                  _lib.interp_code(merge_interp), stream)
    else:
        _lib.call("bpmi_ffbp_merges", _vp(node_int_ptr), _vp(node_geo_ptr), _vp(offsets_ptr),
                  _vp(arena_ptr), _vp(blocks_ptr), int(nblocks), float(k),
                  _lib.mode_code(mode), _lib.interp_code(merge_interp), stream)


class BatchedResidentProgram(ResidentProgram):
    """ResidentProgram-compatible FFBP batched across each tree level."""

    def __init__(self, backend, plan, device, mode, range_start, data_dr, k, chunk_pulses=64,
                 leaf_tile_cols_log2=None, persistent=False, cull=None):
        if persistent and mode != "split":
            raise ValueError("the persistent schedule runs in split precision only")
        if persistent not in (False, True, "merges"):
            raise ValueError("persistent must be False, True or 'merges'")
        if cull is not None and (mode != "split" or persistent == "merges"):
            raise ValueError("culling applies to split-mode schedules with leaves")
        self._cull_request = cull
        self.cull = None
        self.computed_pixels = None  # {node: pixels computed} when culled
        self._table_buffers = []
        self._leaf_tc_request = leaf_tile_cols_log2
        self._batch_stages = []
        self._persistent_request = persistent
        self.persistent = None
        super().__init__(backend, plan, device, mode, range_start, data_dr, k,
                         algorithm="ffbp", variant="fused", chunk_pulses=chunk_pulses)
        self.variant = "batched_fused"
        table_bytes = sum(int(a.nbytes) for a in self._table_buffers)
        self.metadata.update({
            "variant": self.variant,
            "batch_stages": self._batch_stages,
            "extra_table_bytes": table_bytes,
            "extra_table_count": len(self._table_buffers),
            "total_owned_device_bytes": self.metadata["total_owned_device_bytes"] + table_bytes,
            "batch_mapping": "one workgroup per node pixel tile; nodes grouped by subtree height",
            "fusion": "each parent's two child regrids rounded to complex64 then FP32 add",
            "cull": self.cull,
        })

    def _table(self, host):
        array = DeviceArray.from_host(np.ascontiguousarray(host))
        self._table_buffers.append(array)
        return array

    def _build_node(self, root):
        flat = flatten_plan(self.plan)
        self.flat = flat
        arena = self._new_buffer(flat.arena_elems)
        self.arena = arena
        self._node_table = self._table(flat.node_int)
        self._geometry_table = self._table(flat.node_geo)
        self._offset_table = self._table(flat.offsets.astype(np.int64))
        d = self.device
        mode = _lib.mode_code(self.mode)
        tiles = computed = None
        self._cull_tiles = None
        if self.mode == "split":
            tc = self._leaf_tc_request
            if tc is None:
                tc = leaf_tile_width(self.plan, d._host_position, self.data_dr)
            tc = self.leaf_tile_cols_log2 = int(tc)
            if self._cull_request is not None:
                tiles, computed = self._cull(flat, arena, tc)
                self._cull_tiles = tiles
        for height, nodes in enumerate(flat.levels):
            if height == 0:
                if self.mode == "split":
                    mapping = [(n, int(t)) for n in nodes
                               for t in (tiles[n] if tiles is not None else
                                         range(split_tiles(flat.node_int[n, NI_NR],
                                                           flat.node_int[n, NI_NU], tc)))]
                    if not mapping:
                        mapping = [(nodes[0], 0)]
                    from .multipatch import NODE_SOURCE
                    src = np.zeros(flat.node_count, NODE_SOURCE)
                    src["data"], src["pos"] = d.data.ptr, d.pos64.ptr
                    src["weight"], src["phase"] = d.weight64.ptr, d.phase64.ptr
                    src["range_start"], src["data_dr"], src["k"] = (self.range_start,
                                                                    self.data_dr, self.k)
                    src["samples"] = d.data.shape[1]
                    sources = self._table(src.view(np.uint8))
                    blocks = self._table(np.asarray(mapping, dtype=np.int32))
                    self._enqueue("bpmi_ffbp_leaves_multi", len(mapping) * MERGE_BLOCK, (
                        _vp(sources.ptr), _vp(self._node_table.ptr), _vp(self._geometry_table.ptr),
                        _vp(self._offset_table.ptr), _vp(arena.ptr), _vp(blocks.ptr),
                        len(mapping), mode, self._interp(), tc),
                        name=f"batched_bp_polar_{self.mode}")
                else:
                    mapping = [(n, s) for n in nodes
                               for s in range(0, int(flat.node_int[n, NI_NR]) *
                                              int(flat.node_int[n, NI_NU]), MERGE_BLOCK)]
                    blocks = self._table(np.asarray(mapping, dtype=np.int32))
                    self._enqueue("bpmi_ffbp_leaves", len(mapping) * MERGE_BLOCK, (
                        _vp(d.data.ptr), d.data.shape[1], _vp(d.pos64.ptr), _vp(d.weight64.ptr),
                        _vp(d.phase64.ptr), _vp(self._node_table.ptr),
                        _vp(self._geometry_table.ptr), _vp(self._offset_table.ptr),
                        _vp(arena.ptr), _vp(blocks.ptr), len(mapping), self.range_start,
                        self.data_dr, self.k, mode, self._interp()),
                        name=f"batched_bp_polar_{self.mode}")
                kernel = f"batched_bp_polar_{self.mode}"
            else:
                mapping = merge_block_table(flat.node_int, nodes, self.mode, tiles)
                if not mapping:
                    mapping = [(-1, 0)]
                blocks = self._table(np.asarray(mapping, dtype=np.int32))
                if self.mode == "split":
                    self._enqueue("bpmi_ffbp_merges_tiled", len(mapping) * MERGE_BLOCK, (
                        _vp(self._node_table.ptr), _vp(self._geometry_table.ptr),
                        _vp(self._offset_table.ptr), _vp(arena.ptr), _vp(blocks.ptr),
                        len(mapping), self.k, self._merge_interp()),
                        name=f"batched_merge_regrid_{self.mode}")
                else:
                    self._enqueue("bpmi_ffbp_merges", len(mapping) * MERGE_BLOCK, (
                        _vp(self._node_table.ptr), _vp(self._geometry_table.ptr),
                        _vp(self._offset_table.ptr), _vp(arena.ptr), _vp(blocks.ptr),
                        len(mapping), self.k, mode, self._merge_interp()),
                        name=f"batched_merge_regrid_{self.mode}")
                kernel = f"batched_merge_regrid_{self.mode}"
            pixels = int(sum(computed[n] if computed is not None else
                             int(flat.node_int[n, NI_NR]) * int(flat.node_int[n, NI_NU])
                             for n in nodes))
            self.operations[-1].update({"height": height, "node_count": len(nodes),
                                        "valid_pixels": pixels})
            self._batch_stages.append({
                "height": height, "kernel": kernel, "node_count": len(nodes),
                "block_count": len(mapping), "valid_pixels": pixels,
                "block_table_bytes": int(blocks.nbytes)})
        if self._persistent_request:
            self._replace_with_persistent(flat, sources)
        root = flat.offsets[self.plan.root.index]
        # the root image is a view into the arena
        return DeviceArray((self.plan.root.nr * self.plan.root.nu,), np.complex64,
                           arena._alloc, arena._offset + int(root) * 8)

    def _cull(self, flat, arena, tile_cols_log2):
        """Required tiles of every node for the requested output region
        (culling.required_masks); the arena is zeroed once so that pixels
        never computed read as zero."""
        from .culling import required_masks, tile_selection
        masks = required_masks(flat, self.plan, self._node_table, self._geometry_table,
                               region=self._cull_request,
                               merge_interp=self.backend.merge_interp,
                               eval_interp=self.backend.merge_interp)
        tiles, computed = tile_selection(masks, flat, tile_cols_log2)
        self.computed_pixels = computed
        _lib.call("bpmi_zero", _vp(arena.ptr), arena.nbytes, None)
        synchronize()
        total = int(np.sum(flat.node_int[:, NI_NR].astype(np.int64) * flat.node_int[:, NI_NU]))
        self.cull = {
            "region": ("crop" if isinstance(self._cull_request, str) else
                       "cartesian grid %dx%d" % (self._cull_request.ny, self._cull_request.nx)),
            "node_pixels": total,
            "computed_pixels": int(sum(computed.values())),
            "merge_tiles": int(sum(len(tiles[n]) for h in flat.levels[1:] for n in h)),
            "merge_tiles_total": int(sum(merge_tiles(flat.node_int[n, NI_NR],
                                                     flat.node_int[n, NI_NU])
                                         for h in flat.levels[1:] for n in h)),
            "computed_pixels_by_height": [int(sum(computed[n] for n in h)) for h in flat.levels],
            "leaf_tiles": int(sum(len(tiles[n]) for n in flat.levels[0])),
            "leaf_tiles_total": int(sum(split_tiles(flat.node_int[n, NI_NR],
                                                    flat.node_int[n, NI_NU], tile_cols_log2)
                                        for n in flat.levels[0])),
        }
        return tiles, computed

    def _replace_with_persistent(self, flat, sources):
        """Swap the merge launches (persistent=True: and the leaf launch) for
        one persistent launch (bpmi_ffbp_persistent) over the same tables and
        arena."""
        from .persistent import PersistentFFBP
        leaves = self._persistent_request is True
        staged = len(flat.levels) - (0 if leaves else 1)
        del self._schedule[-staged:]
        del self.operations[-staged:]
        self.persistent = PersistentFFBP(
            flat.node_int, self._node_table.ptr, self._geometry_table.ptr, self._offset_table.ptr,
            self.arena.ptr, self.k, self.leaf_tile_cols_log2, sources=sources.ptr,
            leaf_interp=self.backend.interp, merge_interp=self.backend.merge_interp,
            leaves=leaves, tiles=self._cull_tiles)
        pf = self.persistent
        count = pf.nitems * MERGE_BLOCK
        self._enqueue("bpmi_ffbp_persistent", count, (), name=f"ffbp_persistent_{self.mode}",
                      launcher=lambda grid, block, args: pf.launch())
        self.operations[-1].update({"height": None, "node_count": int(flat.node_count),
                                    "work_items": pf.nitems})
        self._batch_stages.append({"height": "all", "kernel": "ffbp_persistent",
                                   "node_count": int(flat.node_count),
                                   "block_count": pf.nitems})

    def node_image(self, index):
        """Device view of node ``index``'s polar image (valid after execute)."""
        n = self.flat.node_int[index]
        off = int(self.flat.offsets[index])
        return DeviceArray((int(n[NI_NR]), int(n[NI_NU])), np.complex64, self.arena._alloc,
                           self.arena._offset + off * 8)

    def close(self):
        super().close()
        if self.persistent is not None:
            self.persistent.close()
        self.persistent = None
        self._table_buffers.clear()
        self._node_table = None
        self._geometry_table = None
        self._offset_table = None
        self.arena = None
