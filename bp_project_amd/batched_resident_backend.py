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
from .resident_backend import ResidentProgram, _Launcher

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


def check_launch_shapes(mode, pulse_block, merge_shapes, merge_heights):
    """Validate a split-mode launch configuration; returns (pulse_block,
    one merge shape per height), defaults filled in."""
    if mode != "split" and (pulse_block is not None or merge_shapes is not None):
        raise ValueError("launch shapes apply to split mode only")
    pb = _lib.DEFAULT_PULSE_BLOCK if pulse_block is None else int(pulse_block)
    if pb not in _lib.PULSE_BLOCKS:
        raise ValueError(f"pulse_block must be one of {sorted(_lib.PULSE_BLOCKS)}")
    if merge_shapes is None:
        shapes = [_lib.DEFAULT_MERGE_SHAPE] * merge_heights
    elif len(merge_shapes) == 2 and all(isinstance(v, (int, np.integer)) for v in merge_shapes):
        shapes = [tuple(int(v) for v in merge_shapes)] * merge_heights
    else:
        shapes = [tuple(int(v) for v in sh) for sh in merge_shapes]
        if len(shapes) != merge_heights:
            raise ValueError(f"merge_shapes needs one shape per merge height ({merge_heights})")
    for sh in shapes:
        if sh not in _lib.MERGE_SHAPES:
            raise ValueError(f"merge shape must be one of {list(_lib.MERGE_SHAPES)}")
    return pb, shapes


def merge_window_table(node_int_ptr, node_geo_ptr, blocks, merge_interp):
    """Window table of a split-mode block table (``bpmi_ffbp_merge_windows``:
    one 64-byte entry per table entry, int32 window bounds (r0, c0, h, w) of
    both children, then their float64 reciprocal spacings; built once, from
    the node geometry only); the tiled merges read it instead of deriving
    the entries in every workgroup."""
    windows = DeviceArray.empty((blocks.shape[0], 16), np.int32)
    _lib.call("bpmi_ffbp_merge_windows", _vp(node_int_ptr), _vp(node_geo_ptr), _vp(blocks.ptr),
              int(blocks.shape[0]), _lib.interp_code(merge_interp), _vp(windows.ptr), None)
    synchronize()
    return windows


def enqueue_merges(node_int_ptr, node_geo_ptr, offsets_ptr, arena_ptr, blocks_ptr, nblocks, k,
                   mode, merge_interp, stream, shape=None, windows_ptr=None):
    """One merge launch over a block table built by ``merge_block_table``;
    ``shape`` = (waves, window_cap) of the split-mode tiled merge,
    ``windows_ptr`` its ``merge_window_table`` (None: derived per tile)."""
    if mode == "split":
        if windows_ptr is not None:
            sh = _lib.DEFAULT_MERGE_SHAPE if shape is None else tuple(shape)
            _lib.call("bpmi_ffbp_merges_tiled_win", _vp(node_int_ptr), _vp(node_geo_ptr),
                      _vp(offsets_ptr), _vp(arena_ptr), _vp(blocks_ptr), int(nblocks), float(k),
                      _lib.interp_code(merge_interp), int(sh[0]), int(sh[1]), _vp(windows_ptr),
                      stream)
            return
        if shape is not None and tuple(shape) != _lib.DEFAULT_MERGE_SHAPE:
            _lib.call("bpmi_ffbp_merges_tiled_ex", _vp(node_int_ptr), _vp(node_geo_ptr),
                      _vp(offsets_ptr), _vp(arena_ptr), _vp(blocks_ptr), int(nblocks), float(k),
                      _lib.interp_code(merge_interp), int(shape[0]), int(shape[1]), stream)
            return
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
                 leaf_tile_cols_log2=None, persistent=False, cull=None, pulse_block=None,
                 merge_shapes=None):
        if persistent and (pulse_block is not None or merge_shapes is not None):
            raise ValueError("launch shapes apply to the level launches, not the persistent "
                             "schedule")
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
        self._shape_request = (pulse_block, merge_shapes)
        self.pulse_block = self.merge_shapes = None
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
        self._leaf_slot, self._merge_slots = None, []
        if self.mode == "split" or self._shape_request != (None, None):
            check_launch_shapes(self.mode, *self._shape_request, len(flat.levels) - 1)
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
                    leaf_args = (
                        _vp(sources.ptr), _vp(self._node_table.ptr), _vp(self._geometry_table.ptr),
                        _vp(self._offset_table.ptr), _vp(arena.ptr), _vp(blocks.ptr),
                        len(mapping), mode, self._interp(), tc)
                    self._enqueue("bpmi_ffbp_leaves_multi", len(mapping) * MERGE_BLOCK,
                                  leaf_args, name=f"batched_bp_polar_{self.mode}")
                    self._leaf_slot = (len(self._schedule) - 1, leaf_args)
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
                    windows = merge_window_table(self._node_table.ptr, self._geometry_table.ptr,
                                                 blocks, self.backend.merge_interp)
                    self._table_buffers.append(windows)
                    merge_args = (
                        _vp(self._node_table.ptr), _vp(self._geometry_table.ptr),
                        _vp(self._offset_table.ptr), _vp(arena.ptr), _vp(blocks.ptr),
                        len(mapping), self.k, self._merge_interp())
                    self._enqueue("bpmi_ffbp_merges_tiled_win", len(mapping) * MERGE_BLOCK,
                                  merge_args + tuple(_lib.DEFAULT_MERGE_SHAPE)
                                  + (_vp(windows.ptr),),
                                  name=f"batched_merge_regrid_{self.mode}")
                    self._merge_slots.append((len(self._schedule) - 1, merge_args,
                                              _vp(windows.ptr)))
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
        elif self.mode == "split":
            self.set_launch_shapes(*self._shape_request)
        root = flat.offsets[self.plan.root.index]
        # the root image is a view into the arena
        return DeviceArray((self.plan.root.nr * self.plan.root.nu,), np.complex64,
                           arena._alloc, arena._offset + int(root) * 8)

    def set_launch_shapes(self, pulse_block=None, merge_shapes=None):
        """Set the split-mode launch shapes of the built schedule in place:
        the leaf kernel's pulse block depth and one (waves, window_cap)
        merge shape per subtree height (a single pair applies to every
        height; None = the defaults).  A captured graph is dropped (the
        caller first waits for work that still uses it)."""
        if self.persistent is not None or self._leaf_slot is None:
            raise ValueError("launch shapes apply to split-mode level launches")
        pb, shapes = check_launch_shapes(self.mode, pulse_block, merge_shapes,
                                         len(self._merge_slots))
        index, args = self._leaf_slot
        if pb == _lib.DEFAULT_PULSE_BLOCK:
            self._set_entry(index, "bpmi_ffbp_leaves_multi", args)
        else:
            self._set_entry(index, "bpmi_ffbp_leaves_multi_ex", args + (pb,))
        for (index, args, windows), shape in zip(self._merge_slots, shapes):
            self._set_entry(index, "bpmi_ffbp_merges_tiled_win", args + tuple(shape) + (windows,))
        self.pulse_block, self.merge_shapes = pb, shapes
        if self._graph is not None:
            self._graph.close()
            self._graph = None
        if self._graph_stream is not None:
            self._graph_stream.close()
            self._graph_stream = None

    def _set_entry(self, index, cname, args):
        label, _, grid, block, _ = self._schedule[index]
        launcher = _Launcher(cname)
        self._schedule[index] = (label, launcher, grid, block, args)
        self.operations[index].update({"kernel": launcher, "args": args})

    def launch_stage_names(self):
        """Stage names of the schedule's launches as ``autotune.stage_times``
        reports them when every launch is marked with its label."""
        from .autotune import stage_keys
        return stage_keys([e[0] for e in self._schedule])

    def autotune(self, execute=None, force=False, pulse_blocks=None, merge_shapes=None,
                 warmup=1, repeats=3):
        """Choose the leaf pulse block depth and one merge shape per subtree
        height by timing (autotune.tune_config); the leaf tile width stays
        as built.  ``execute(mark)`` runs one step and calls ``mark(label)``
        after each launch (default: the schedule, marked with its labels;
        a caller may add launches such as an evaluation).  Returns (config,
        report, cached) and leaves the schedule on the chosen shapes."""
        from .autotune import Group, make_key, stage_keys, stage_times, tune_config
        if self.persistent is not None or self._leaf_slot is None:
            raise ValueError("launch shapes apply to split-mode level launches")
        if execute is None:
            def execute(mark):
                for label, launcher, grid, block, args in self._schedule:
                    launcher(grid, block, args)
                    if mark:
                        mark(label)
        names = stage_keys([e[0] for e in self._schedule])
        leaf_stage = names[self._leaf_slot[0]]
        merge_stages = [names[i] for i, _, _ in self._merge_slots]
        groups = [Group("pulse_block", list(pulse_blocks or _lib.PULSE_BLOCKS), [leaf_stage]),
                  Group("merge_shapes", [list(s) for s in (merge_shapes or _lib.MERGE_SHAPES)],
                        merge_stages, per_stage=True)]
        work = int(self.plan.counts["leaf_pixel_pulse_products"])
        key = make_key("ffbp_levels", self.mode,
                       f"{self.backend.interp}+{self.backend.merge_interp}", work,
                       extra=f"h{len(merge_stages)}/tc{self.leaf_tile_cols_log2}"
                             f"/{'cull' if self.cull else 'full'}")

        def measure(config):
            self.set_launch_shapes(config["pulse_block"], config["merge_shapes"])
            synchronize()
            return stage_times(execute, warmup, repeats)
        base = {"pulse_block": self.pulse_block,
                "merge_shapes": [list(s) for s in self.merge_shapes]}
        if not groups[0].valid(base["pulse_block"]):
            base["pulse_block"] = groups[0].candidates[0]
        if not groups[1].valid(base["merge_shapes"]):
            base["merge_shapes"] = [groups[1].candidates[0]] * len(merge_stages)
        config, report, cached = tune_config(key, groups, measure, base, force=force)
        self.set_launch_shapes(config["pulse_block"], config["merge_shapes"])
        return config, report, cached

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
