"""Batched FFBP serving of many independent, uneven patches.

All leaves of all patches run in one launch (per-node input table), the
merges of every patch run one launch per subtree height, and every patch's
Cartesian output is evaluated in one launch (split mode: LDS-tiled, 64 x 4
pixel tiles); the whole batch is captured
in one HIP graph, so a batch is one graph launch with no host round trips.
In split mode the launch shapes (leaf tile width, the leaf kernel's pulse
block depth, one merge shape per subtree height) are chosen by the
autotuner.  With
``cull=True`` only the node tiles that each patch's Cartesian output depends
on are launched (culling.py); the outputs are bitwise unchanged.

Counterpart of nothing in the reference (it images one ROI at a time,
fast_bp/batched_resident_backend.py); it composes the reference-parity
kernels: results are bitwise equal to running each patch's
``BatchedResidentProgram`` + evaluation alone with the same tile width.
"""
from __future__ import annotations

import json

import numpy as np

from . import _lib
from .device import DeviceArray, Event, Graph, Stream, current_stream, elapsed_ms, stream_handle, synchronize
from .batched_resident_backend import (check_launch_shapes, enqueue_merges, merge_block_table,
                                       merge_window_table)
from .culling import merge_tiles_of, split_tiles_of
from .gpu_backend import build_ffbp_plan, ensure_kaiser_table
from .plan import NI_HEIGHT, NI_LEFT, NI_NR, NI_NU, NI_RIGHT, flatten_plan
from .scenes import CartesianGrid, polar_extent_for_grid, stripmap_scene

_vp = _lib.c_void_p
BLOCK = 256

NODE_SOURCE = np.dtype([("data", "<u8"), ("pos", "<u8"), ("weight", "<u8"), ("phase", "<u8"),
                        ("range_start", "<f8"), ("data_dr", "<f8"), ("k", "<f8"),
                        ("samples", "<i4"), ("reserved", "<i4")])
EVAL_JOB = np.dtype([("img", "<u8"), ("out", "<u8"), ("geo", "<f8", (7,)), ("x0", "<f8"),
                     ("dx", "<f8"), ("y0", "<f8"), ("dy", "<f8"), ("k", "<f8"), ("nr", "<i4"),
                     ("nu", "<i4"), ("ld", "<i4"), ("nx", "<i4"), ("ny", "<i4"),
                     ("reserved", "<i4")])
assert NODE_SOURCE.itemsize == 64 and EVAL_JOB.itemsize == 136


def patch_mix(count=64, seed=20260921, scale=1.0):
    """Seeded mix of uneven patches: pulses from {256..4096}, bins from
    {1024, 2048, 4096}, square grids of 128..1024 pixels at 0.5 m, 10..100
    targets, 20 dB clutter.  ``scale`` < 1 shrinks every size (tests)."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        pulses = int(rng.choice([256, 512, 1024, 2048, 4096]))
        bins = int(rng.choice([1024, 2048, 4096]))
        grid_n = int(rng.choice([128, 256, 384, 512, 640, 768, 896, 1024]))
        targets = int(rng.integers(10, 101))
        if scale != 1.0:
            pulses = max(32, int(pulses * scale))
            bins = max(256, int(bins * scale))
            grid_n = max(16, int(grid_n * scale))
            targets = max(2, int(targets * scale))
        # a bin count must cover the grid's slant extent
        bins = max(bins, int(grid_n * 0.5 * 1.3 / (299792458.0 / (2 * 180e6))) + 256)
        specs.append(dict(pulses=pulses, bins=bins, n_targets=targets, grid_n=grid_n,
                          grid_spacing=0.5, seed=seed * 1000 + i))
    return specs


def make_patches(specs):
    """The scenes of ``specs``, generated concurrently (one thread per
    scene; each scene is the same as generated alone)."""
    from concurrent.futures import ThreadPoolExecutor
    from .scenes import _scene_threads
    with ThreadPoolExecutor(max_workers=_scene_threads()) as pool:
        return list(pool.map(lambda s: stripmap_scene(**s, threads=1), specs))


class PatchPlan:
    def __init__(self, scene, leaf_size, range_spacing, angular_oversample):
        self.scene = scene
        extent = polar_extent_for_grid(scene.grid, scene.pos.mean(axis=0), margin=2.0)
        self.plan = build_ffbp_plan(scene.pos, scene.fc, extent, side=1, leaf_size=leaf_size,
                                    range_spacing=range_spacing,
                                    angular_oversample=angular_oversample)
        self.flat = flatten_plan(self.plan)


class MultiPatchProgram:
    """Level-batched FFBP + Cartesian evaluation of many patches."""

    def __init__(self, scenes, mode="split", leaf_size=16, range_spacing=0.5,
                 angular_oversample=2.5, leaf_interp="lagrange4", merge_interp="kaiser8",
                 tile_cols_log2=6, persistent=False, cull=False, pulse_block=None,
                 merge_shapes=None):
        if not scenes:
            raise ValueError("no patches")
        if persistent and (pulse_block is not None or merge_shapes is not None):
            raise ValueError("launch shapes apply to the level launches, not the persistent "
                             "schedule")
        if cull and (mode != "split" or persistent == "merges"):
            raise ValueError("culling applies to split-mode schedules with leaves")
        ks = {round(s.k, 9) for s in scenes}
        if len(ks) != 1:
            raise ValueError("all patches of a batch must share the carrier wavenumber")
        if "kaiser8" in (leaf_interp, merge_interp):
            ensure_kaiser_table()
        if persistent and mode != "split":
            raise ValueError("the persistent schedule runs in split precision only")
        if persistent not in (False, True, "merges"):
            raise ValueError("persistent must be False, True or 'merges'")
        self.persistent = persistent
        self._persistent = None
        self.mode = mode
        self.leaf_interp = leaf_interp
        self.merge_interp = merge_interp
        self.k = float(scenes[0].k)
        self.patches = [PatchPlan(s, leaf_size, range_spacing, angular_oversample) for s in scenes]
        self._inputs = []
        node_int, node_geo, sizes, sources = [], [], [], []
        self.node_base = []
        base = 0
        for pp in self.patches:
            s = pp.scene
            d = DeviceArray.from_host(s.data)
            p = DeviceArray.from_host(s.pos)
            w = DeviceArray.from_host(np.ascontiguousarray(s.weight, np.float64))
            ph = DeviceArray.from_host(np.ascontiguousarray(s.phase, np.float64))
            self._inputs += [d, p, w, ph]
            ni = pp.flat.node_int.copy()
            for col in (NI_LEFT, NI_RIGHT):
                ni[:, col] = np.where(ni[:, col] >= 0, ni[:, col] + base, -1)
            node_int.append(ni)
            node_geo.append(pp.flat.node_geo)
            sizes.append(ni[:, NI_NR].astype(np.int64) * ni[:, NI_NU])
            src = np.zeros(len(ni), NODE_SOURCE)
            src["data"], src["pos"], src["weight"], src["phase"] = d.ptr, p.ptr, w.ptr, ph.ptr
            src["range_start"], src["data_dr"], src["k"] = s.r_start, s.data_dr, s.k
            src["samples"] = s.bins
            sources.append(src)
            self.node_base.append(base)
            base += len(ni)
        self.node_int = np.ascontiguousarray(np.concatenate(node_int), np.int32)
        self.node_geo = np.ascontiguousarray(np.concatenate(node_geo))
        sizes = np.concatenate(sizes)
        offsets = np.zeros(len(sizes), np.int64)
        cursor = 0
        for i, n in enumerate(sizes):
            offsets[i] = cursor
            cursor += -(-int(n) // 64) * 64
        self.offsets = offsets
        self.arena = DeviceArray.empty(cursor)
        self.t_node_int = DeviceArray.from_host(self.node_int)
        self.t_node_geo = DeviceArray.from_host(self.node_geo)
        self.t_offsets = DeviceArray.from_host(offsets)
        self.t_sources = DeviceArray.from_host(np.concatenate(sources).view(np.uint8))
        self._masks = self._required_masks() if cull else None
        heights = self.node_int[:, NI_HEIGHT]
        self.pulse_block, self.merge_shapes = (
            check_launch_shapes(mode, pulse_block, merge_shapes, int(heights.max()))
            if mode == "split" or pulse_block is not None or merge_shapes is not None
            else (None, None))
        self.set_tile_width(tile_cols_log2)
        self.merge_tables = []
        self.merge_windows = []  # split mode: the window bounds of each table
        tiles = (None if self._masks is None else
                 {n: merge_tiles_of(m) for n, m in enumerate(self._masks)})
        self.merge_pixels = []  # parent pixels each merge launch computes
        for h in range(1, int(heights.max()) + 1):
            nodes = np.flatnonzero(heights == h)
            mapping = merge_block_table(self.node_int, nodes, self.mode, tiles) or [(-1, 0)]
            self.merge_tables.append(DeviceArray.from_host(np.asarray(mapping, np.int32)))
            if self.mode == "split":
                self.merge_windows.append(merge_window_table(
                    self.t_node_int.ptr, self.t_node_geo.ptr, self.merge_tables[-1],
                    self.merge_interp))
            self.merge_pixels.append(self._mapping_pixels(mapping))
        # Cartesian evaluation jobs (crop as a view into each root)
        jobs = np.zeros(len(self.patches), EVAL_JOB)
        self.outputs = []
        eval_blocks = []
        for j, pp in enumerate(self.patches):
            plan, g = pp.plan, pp.scene.grid
            root = self.node_base[j] + plan.root.index
            out = DeviceArray.empty(g.shape)
            self.outputs.append(out)
            img = self.arena.ptr + 8 * (int(offsets[root]) + plan.crop_r[0] * plan.root.nu
                                        + plan.crop_u[0])
            jobs[j]["img"], jobs[j]["out"] = img, out.ptr
            jobs[j]["geo"] = (plan.origin[0], plan.origin[1], plan.origin[2], plan.r0, plan.dr,
                              plan.u0, plan.du)
            jobs[j]["x0"], jobs[j]["dx"], jobs[j]["y0"], jobs[j]["dy"] = g.x0, g.dx, g.y0, g.dy
            jobs[j]["k"] = self.k
            jobs[j]["nr"], jobs[j]["nu"], jobs[j]["ld"] = plan.nr, plan.nu, plan.root.nu
            jobs[j]["nx"], jobs[j]["ny"] = g.nx, g.ny
            if self.mode == "split":
                import ctypes
                t = ctypes.c_int()
                _lib.call("bpmi_eval_tiles", g.nx, g.ny, ctypes.byref(t))
                eval_blocks += [(j, i) for i in range(t.value)]
            else:
                eval_blocks += [(j, s) for s in range(0, g.nx * g.ny, BLOCK)]
        self.jobs_host = jobs
        self.job_roots = [self.node_base[j] + pp.plan.root.index
                          for j, pp in enumerate(self.patches)]
        self.t_jobs = DeviceArray.from_host(jobs.view(np.uint8))
        self.t_eval_blocks = DeviceArray.from_host(np.asarray(eval_blocks, np.int32))
        self.n_eval_blocks = len(eval_blocks)
        self._graph = None
        self._stream = None
        self.launch_count = 2 + len(self.merge_tables)
        if self.persistent:
            self._build_persistent()
            self.launch_count = 1 if self.persistent is True else 2

    def _mapping_pixels(self, mapping):
        """Parent pixels covered by a merge block table (32 x 32 tiles clipped
        to their node in split mode, else blocks of merge pixels)."""
        from .batched_resident_backend import MERGE_BLOCK
        total = 0
        for n, t in mapping:
            if n < 0:
                continue
            nr, nu = int(self.node_int[n, NI_NR]), int(self.node_int[n, NI_NU])
            if self.mode == "split":
                tu = -(-nu // 32)
                r0, c0 = (t // tu) * 32, (t % tu) * 32
                total += (min(r0 + 32, nr) - r0) * (min(c0 + 32, nu) - c0)
            else:
                total += min(MERGE_BLOCK, nr * nu - t)
        return total

    def _required_masks(self):
        """Tile masks of every node (global index) that each patch's
        Cartesian output depends on (culling.required_masks per patch); the
        arena is zeroed once so that tiles never computed read as zero."""
        from .culling import required_masks
        masks = []
        for pp in self.patches:
            nt = DeviceArray.from_host(pp.flat.node_int)
            gt = DeviceArray.from_host(pp.flat.node_geo)
            masks += required_masks(pp.flat, pp.plan, nt, gt, region=pp.scene.grid,
                                    merge_interp=self.merge_interp,
                                    eval_interp=self.merge_interp)
        _lib.call("bpmi_zero", _vp(self.arena.ptr), self.arena.nbytes, None)
        synchronize()
        return masks

    def _selected_tiles(self):
        """{node: tile indices} of the culled schedule (None without culling)."""
        if self._masks is None:
            return None
        heights = self.node_int[:, NI_HEIGHT]
        return {n: (split_tiles_of(m, int(self.node_int[n, NI_NR]), int(self.node_int[n, NI_NU]),
                                   self.tile_cols_log2) if heights[n] == 0 else merge_tiles_of(m))
                for n, m in enumerate(self._masks)}

    def set_tile_width(self, tile_cols_log2):
        """(Re)build the leaf block table for a split-mode tile width."""
        self.tile_cols_log2 = int(tile_cols_log2)
        leaves = np.flatnonzero(self.node_int[:, NI_HEIGHT] == 0)
        mapping = []
        if self.mode == "split" and self._masks is not None:
            for n in leaves:
                sel = split_tiles_of(self._masks[n], int(self.node_int[n, NI_NR]),
                                     int(self.node_int[n, NI_NU]), self.tile_cols_log2)
                mapping += [(int(n), int(i)) for i in sel]
        elif self.mode == "split":
            import ctypes
            t = ctypes.c_int()
            for n in leaves:
                _lib.call("bpmi_split_tiles_ex", int(self.node_int[n, NI_NR]),
                          int(self.node_int[n, NI_NU]), self.tile_cols_log2, ctypes.byref(t))
                mapping += [(int(n), i) for i in range(t.value)]
        else:
            for n in leaves:
                size = int(self.node_int[n, NI_NR]) * int(self.node_int[n, NI_NU])
                mapping += [(int(n), s) for s in range(0, size, BLOCK)]
        self.leaf_blocks = DeviceArray.from_host(np.asarray(mapping, np.int32))
        self.n_leaf_blocks = len(mapping)
        if getattr(self, "_graph", None) is not None:
            self._graph.close()
            self._graph = None
        if getattr(self, "_persistent", None) is not None:
            self._build_persistent()

    def _build_persistent(self):
        """One persistent launch for the whole batch.  Within each height,
        nodes of patches with fewer levels go first, so those patches
        complete early."""
        from .persistent import PersistentFFBP
        root_height = np.zeros(len(self.node_int), np.int64)
        for j, pp in enumerate(self.patches):
            n0 = self.node_base[j]
            root_height[n0:n0 + len(pp.flat.node_int)] = pp.flat.node_int[pp.plan.root.index,
                                                                          NI_HEIGHT]
        if self._persistent is not None:
            self._persistent.close()
        self._persistent = PersistentFFBP(
            self.node_int, self.t_node_int.ptr, self.t_node_geo.ptr, self.t_offsets.ptr,
            self.arena.ptr, self.k, self.tile_cols_log2, sources=self.t_sources.ptr,
            jobs=self.jobs_host, job_roots=self.job_roots, leaf_interp=self.leaf_interp,
            merge_interp=self.merge_interp, priority=root_height,
            leaves=self.persistent is True, tiles=self._selected_tiles())

    def completion_ms(self):
        """Persistent schedule, after a completed launch: each patch's
        completion time in ms after the launch's first workgroup started."""
        if self._persistent is None:
            raise RuntimeError("completion times need persistent=True")
        if self._persistent.aborted():
            raise RuntimeError("persistent FFBP aborted: a dependency wait timed out")
        return self._persistent.completion_ms()

    @property
    def work(self):
        """(output pixels x pulses) summed over patches."""
        return sum(pp.scene.grid.nx * pp.scene.grid.ny * pp.scene.pulses for pp in self.patches)

    def execute(self, mark=None):
        """Enqueue the batch on the current stream.  ``mark(name)``, when
        given, is called after each launch ("leaves", "merges", "eval") so a
        caller can record events between them."""
        mode = _lib.mode_code(self.mode)
        s = stream_handle()
        if self._persistent is not None:
            if not self._persistent.leaves:
                self._launch_leaves(mode, s)
            self._persistent.launch(s)
            return self.outputs
        self._launch_leaves(mode, s)
        if mark:
            mark("leaves")
        for h, table in enumerate(self.merge_tables):
            enqueue_merges(self.t_node_int.ptr, self.t_node_geo.ptr, self.t_offsets.ptr,
                           self.arena.ptr, table.ptr, table.shape[0], self.k, self.mode,
                           self.merge_interp, s,
                           shape=self.merge_shapes[h] if self.merge_shapes else None,
                           windows_ptr=self.merge_windows[h].ptr if self.merge_windows else None)
            if mark:
                mark("merges")
        if self.mode == "split":
            _lib.call("bpmi_eval_grid_multi_tiled", _vp(self.t_jobs.ptr),
                      _vp(self.t_eval_blocks.ptr), self.n_eval_blocks,
                      _lib.interp_code(self.merge_interp), s)
        else:
            _lib.call("bpmi_eval_grid_multi", _vp(self.t_jobs.ptr), _vp(self.t_eval_blocks.ptr),
                      self.n_eval_blocks, mode, _lib.interp_code(self.merge_interp), s)
        if mark:
            mark("eval")
        return self.outputs

    def _launch_leaves(self, mode, s):
        args = (_vp(self.t_sources.ptr), _vp(self.t_node_int.ptr), _vp(self.t_node_geo.ptr),
                _vp(self.t_offsets.ptr), _vp(self.arena.ptr), _vp(self.leaf_blocks.ptr),
                self.n_leaf_blocks, mode, _lib.interp_code(self.leaf_interp), self.tile_cols_log2)
        if self.pulse_block not in (None, _lib.DEFAULT_PULSE_BLOCK):
            _lib.call("bpmi_ffbp_leaves_multi_ex", *args, self.pulse_block, s)
        else:
            _lib.call("bpmi_ffbp_leaves_multi", *args, s)

    def capture_graph(self):
        if self._graph is None:
            self._stream = Stream()
            current_stream().synchronize()
            with self._stream:
                self.execute()
            self._stream.synchronize()
            self._graph = Graph.capture(self._stream, self.execute)
        return True

    def replay(self):
        if self._graph is None:
            raise RuntimeError("capture_graph() must succeed before replay()")
        self._graph.launch()
        return self.outputs

    def launch_config(self):
        """The split-mode launch shapes in use: {"leaf": [tile_cols_log2,
        pulse_block], "merge_shapes": [[waves, window_cap] per height]}."""
        return {"leaf": [self.tile_cols_log2, self.pulse_block],
                "merge_shapes": [list(sh) for sh in self.merge_shapes]}

    def set_launch_config(self, config):
        """Apply a ``launch_config()``-shaped dict (level launches, split
        mode); rebuilds the leaf table when the tile width changes and drops
        a captured graph."""
        if self.mode != "split" or self._persistent is not None:
            raise ValueError("launch configurations apply to split-mode level launches")
        tc, pb = (int(v) for v in config["leaf"])
        pb, shapes = check_launch_shapes(self.mode, pb, config["merge_shapes"],
                                         len(self.merge_tables))
        changed = (pb, shapes) != (self.pulse_block, self.merge_shapes)
        self.pulse_block, self.merge_shapes = pb, shapes
        if tc != self.tile_cols_log2:
            self.set_tile_width(tc)
        elif changed and self._graph is not None:
            self._graph.close()
            self._graph = None

    def autotune(self, force=False, tile_widths=(4, 5, 6), pulse_blocks=None,
                 merge_shapes=None, warmup=1, repeats=3):
        """Choose the split-mode launch shapes by timing (autotune.tune_config).

        Two knob groups, each timed per launch stage with the other at its
        current best: the leaf launch's (tile width, pulse block depth)
        pairs, then the merge shapes, one chosen per subtree height from the
        merge launch of that height.  Returns (config, report, cached) and
        leaves the program on the chosen config.  A persistent schedule
        tunes its leaf tile width only (one launch: whole-batch time).
        """
        from .autotune import Group, make_key, tune, tune_config
        if self.mode != "split":
            return None, {}, False
        work = int(sum(pp.plan.counts["leaf_pixel_pulse_products"] for pp in self.patches))
        if self._persistent is not None:
            key = make_key("ffbp_persistent_multi", self.mode, self.leaf_interp, work)

            def run(tc):
                # the rebuild happens in the untimed warm-up call of a candidate
                if tc != self.tile_cols_log2:
                    self.set_tile_width(tc)
                    synchronize()
                self.execute()
            best, times = tune(key, list(tile_widths), run, force=force)
            self.set_tile_width(best)
            # the report in tune_config's shape: one stage (the whole launch)
            report = {"leaf": {"stages": ["batch"], "choice": [best, None],
                               "times": {json.dumps([int(k), None]): [float(v)]
                                         for k, v in times.items()}}}
            return {"leaf": [best, None]}, report, False
        pulse_blocks = tuple(pulse_blocks or _lib.PULSE_BLOCKS)
        merge_shapes = tuple(merge_shapes or _lib.MERGE_SHAPES)
        heights = len(self.merge_tables)
        merge_stages = [f"merges{h}" for h in range(1, heights + 1)] if heights > 1 else ["merges"]
        groups = [Group("leaf", [[tc, pb] for tc in tile_widths for pb in pulse_blocks],
                        ["leaves"]),
                  Group("merge_shapes", [list(sh) for sh in merge_shapes], merge_stages,
                        per_stage=True)]
        key = make_key("multipatch_levels", self.mode,
                       f"{self.leaf_interp}+{self.merge_interp}", work,
                       extra=f"h{heights}/{'cull' if self._masks is not None else 'full'}")

        def measure(config):
            from .autotune import stage_times
            self.set_launch_config(config)
            synchronize()
            return stage_times(self.execute, warmup, repeats)
        base = self.launch_config()
        if not groups[0].valid(base["leaf"]):
            base["leaf"] = groups[0].candidates[0]
        config, report, cached = tune_config(key, groups, measure, base, force=force)
        self.set_launch_config(config)
        return config, report, cached

    def close(self):
        if self._graph is not None:
            self._graph.close()
        self._graph = None
        if self._persistent is not None:
            self._persistent.close()
        self._persistent = None
        self.outputs = []
        self._inputs = []
        self.arena = None


def group_by_work(work, fractions):
    """Indices 0..len(work)-1 ordered by work (ties by index) and cut at the
    given fractions of the count (rounded; empty groups dropped)."""
    n = len(work)
    order = sorted(range(n), key=lambda i: (work[i], i))
    cuts = sorted({int(round(f * n)) for f in fractions if 0.0 < f < 1.0})
    bounds = [0] + [c for c in cuts if 0 < c < n] + [n]
    return [order[a:b] for a, b in zip(bounds, bounds[1:])]


class GroupedMultiPatch:
    """A batch served in latency groups.

    The patches are sorted by work (output pixels x pulses) and split at the
    given fractions of the patch count into consecutive groups, each a
    MultiPatchProgram; the groups run one after another (one HIP graph holds
    them all).  A patch completes with its group's evaluation launch, so the
    small patches of the first groups complete after a fraction of the batch
    time, at the price of a less filled GPU while those groups run.  Outputs
    are in the order of ``scenes`` and are the same arrays as one
    MultiPatchProgram of all patches computes with the same leaf shapes (a
    patch's images do not depend on the other patches of its launch).
    Keyword arguments go to every group's MultiPatchProgram.
    """

    def __init__(self, scenes, fractions=(0.5, 0.75), **kwargs):
        if not scenes:
            raise ValueError("no patches")
        n = len(scenes)
        self.groups = group_by_work([s.grid.nx * s.grid.ny * s.pulses for s in scenes], fractions)
        self.programs = [MultiPatchProgram([scenes[i] for i in g], **kwargs) for g in self.groups]
        self.outputs = [None] * n
        for g, prog in zip(self.groups, self.programs):
            for i, o in zip(g, prog.outputs):
                self.outputs[i] = o
        self.launch_count = sum(p.launch_count for p in self.programs)
        self._graph = None
        self._stream = None

    @property
    def work(self):
        return sum(p.work for p in self.programs)

    def execute(self, group_done=None):
        """Enqueue every group on the current stream; ``group_done(g)``, when
        given, is called after group g's launches."""
        for g, prog in enumerate(self.programs):
            prog.execute()
            if group_done:
                group_done(g)
        return self.outputs

    def autotune(self, force=False, leaf=None):
        """Tune each group's launch shapes (MultiPatchProgram.autotune);
        ``leaf`` = (tile width, pulse block) fixes the leaf shapes, e.g. to
        those of a batch whose outputs these must equal bit for bit.  Returns
        the groups' (config, report, cached)."""
        out = []
        for prog in self.programs:
            if leaf is None:
                out.append(prog.autotune(force=force))
            else:
                pb = _lib.DEFAULT_PULSE_BLOCK if leaf[1] is None else int(leaf[1])
                out.append(prog.autotune(force=force, tile_widths=(int(leaf[0]),),
                                         pulse_blocks=(pb,)))
        if self._graph is not None:
            self._graph.close()
            self._graph = None
        return out

    def completion_ms(self):
        """One eager execution: each patch's completion time (ms after the
        batch started, HIP events after each group), in ``scenes`` order."""
        start = Event()
        ends = []

        def group_done(g):
            e = Event()
            e.record()
            ends.append(e)
        start.record()
        self.execute(group_done)
        ends[-1].synchronize()
        done = np.zeros(len(self.outputs))
        for g, e in zip(self.groups, ends):
            done[g] = elapsed_ms(start, e)
        return done

    def capture_graph(self):
        if self._graph is None:
            self._stream = Stream()
            current_stream().synchronize()
            with self._stream:
                self.execute()
            self._stream.synchronize()
            self._graph = Graph.capture(self._stream, self.execute)
        return True

    def replay(self):
        if self._graph is None:
            raise RuntimeError("capture_graph() must succeed before replay()")
        self._graph.launch()
        return self.outputs

    def close(self):
        if self._graph is not None:
            self._graph.close()
        self._graph = None
        for prog in self.programs:
            prog.close()
        self.programs = []
        self.outputs = []


def sequential_reference(scenes, mode="split", leaf_size=16, range_spacing=0.5,
                         angular_oversample=2.5, leaf_interp="lagrange4", merge_interp="kaiser8",
                         tile_cols_log2=6, cull=False, pulse_block=None):
    """Per-patch programs (the same kernels, one patch at a time): returns the
    outputs, each patch's isolated event time and its completion time when
    the patches are launched back to back.  ``pulse_block`` is the leaf
    kernel's pulse block depth (the merge shapes never change a result)."""
    from .batched_resident_backend import BatchedResidentProgram
    from .gpu_backend import HipBaseline, geo_array
    backend = HipBaseline(interp=leaf_interp, merge_interp=merge_interp)
    progs = []
    for s in scenes:
        pp = PatchPlan(s, leaf_size, range_spacing, angular_oversample)
        dev = backend.upload(s.data, s.pos, s.weight, s.phase, pp.plan)
        prog = BatchedResidentProgram(backend, pp.plan, dev, mode, s.r_start, s.data_dr, s.k,
                                      leaf_tile_cols_log2=tile_cols_log2,
                                      cull=s.grid if cull else None, pulse_block=pulse_block)
        out = DeviceArray.empty(s.grid.shape)
        progs.append((pp, prog, out, geo_array(pp.plan)))

    def run_one(entry):
        pp, prog, out, geo = entry
        g = pp.scene.grid
        prog.execute()
        _lib.call("bpmi_eval_grid", _vp(prog.output.ptr), geo.ctypes.data_as(_vp), pp.plan.nr,
                  pp.plan.nu, pp.plan.nu, g.x0, g.dx, g.nx, g.y0, g.dy, g.ny, pp.scene.k,
                  _lib.mode_code(mode), _lib.interp_code(merge_interp), _vp(out.ptr),
                  stream_handle())
    for e in progs:
        run_one(e)
    synchronize()
    start = Event()
    start.record()
    marks = []
    for e in progs:
        a, b = Event(), Event()
        a.record()
        run_one(e)
        b.record()
        marks.append((a, b))
    synchronize()
    isolated = [elapsed_ms(a, b) for a, b in marks]
    completion = [elapsed_ms(start, b) for _, b in marks]
    outputs = [e[2] for e in progs]
    return outputs, isolated, completion, progs
