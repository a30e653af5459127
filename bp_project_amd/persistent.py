"""One-launch FFBP: the work list and workspace of ``bpmi_ffbp_persistent``.

The level-by-level programs (``BatchedResidentProgram``, ``MultiPatchProgram``)
launch the leaves, one merge level per subtree height and the Cartesian
evaluation as separate kernels.  Here the same items -- leaf split tiles,
32 x 32 merge tiles, 64 x 4-pixel evaluation tiles (the LDS-tiled evaluation
of the level programs) -- go into one list that a
single persistent launch drains: workgroups claim items in list order, a
merge tile waits until both of its children are complete, an evaluation
block until its root is.  The list is height-major (all leaves, then every
height-1 merge, ...), and the evaluation tiles of a patch follow the merges
of its root's height, so small patches finish early while the GPU keeps
working on the upper levels of large ones.  Within a height the nodes are
taken in ``priority`` order (default: as given).

Each job's completion time (device wall clock, 100 MHz) is stamped by the
kernel, which gives per-patch latency inside one launch.  Outputs are
bitwise equal to the level-by-level launches (same tile bodies, same
arithmetic).  Split precision only.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .device import DeviceArray, stream_handle
from .plan import NI_HEIGHT, NI_NR, NI_NU

WORK_LEAF, WORK_MERGE, WORK_EVAL = 0, 1, 2
INTERP_NONE = -1  # BPMI_INTERP_NONE: a work list without leaf items



def _split_tiles(nr, nu, tc):
    t = ctypes.c_int()
    _lib.call("bpmi_split_tiles_ex", int(nr), int(nu), int(tc), ctypes.byref(t))
    return t.value


def _merge_tiles(nr, nu):
    t = ctypes.c_int()
    _lib.call("bpmi_merge_tiles", int(nr), int(nu), ctypes.byref(t))
    return t.value


def _eval_tiles(nx, ny):
    t = ctypes.c_int()
    _lib.call("bpmi_eval_tiles", int(nx), int(ny), ctypes.byref(t))
    return t.value


def work_items(node_int, tile_cols_log2, jobs=(), priority=None, leaves=True, tiles=None):
    """Work list of a forest of FFBP trees.

    node_int: int32 [nodes, 8] node table (a, b, nr, nu, side, left, right,
    height); jobs: sequence of (root node, nx, ny) per evaluation job (an
    nx x ny Cartesian grid, evaluated in bpmi_eval_tiles() tiles);
    leaves=False leaves the leaf tiles out (they are computed before the
    launch; their completion targets are then 0); tiles: optional {node:
    tile indices} of a culled schedule (culling.tile_selection), every node's
    completion target then being its number of listed tiles.
    Returns (items int32 [n, 4], node_tiles int32 [nodes], job_blocks int32
    [jobs]).  Every item comes after the items it depends on.
    """
    node_int = np.asarray(node_int)
    nodes = len(node_int)
    heights = node_int[:, NI_HEIGHT].astype(np.int64)
    prio = np.arange(nodes) if priority is None else np.asarray(priority)
    node_tiles = np.zeros(nodes, np.int32)
    tile_list = {}
    for n in range(nodes):
        nr, nu = int(node_int[n, NI_NR]), int(node_int[n, NI_NU])
        if tiles is not None:
            tile_list[n] = [int(t) for t in tiles[n]]
        elif heights[n] == 0:
            tile_list[n] = list(range(_split_tiles(nr, nu, tile_cols_log2)))
        else:
            tile_list[n] = list(range(_merge_tiles(nr, nu)))
        if heights[n] == 0 and not leaves:
            tile_list[n] = []
        node_tiles[n] = len(tile_list[n])
    job_blocks = np.array([_eval_tiles(nx, ny) for _, nx, ny in jobs], np.int32)
    job_height = [int(heights[root]) for root, _, _ in jobs]
    items = []
    for h in range(int(heights.max()) + 1 if nodes else 0):
        level = sorted(np.flatnonzero(heights == h), key=lambda n: (prio[n], n))
        kind = WORK_LEAF if h == 0 else WORK_MERGE
        for n in level:
            items.extend((kind, int(n), t, 0) for t in tile_list[int(n)])
        for j in range(len(jobs)):
            if job_height[j] == h:
                items.extend((WORK_EVAL, j, b, 0) for b in range(int(job_blocks[j])))
    return np.asarray(items, np.int32).reshape(-1, 4), node_tiles, job_blocks


def check_order(items, node_int, job_root, leaves=True):
    """True when every item follows all items it depends on (host check of a
    work list; the kernel's deadlock freedom rests on it).  leaves=False:
    leaf nodes are complete before the launch."""
    node_int = np.asarray(node_int)
    last = {}
    if not leaves:
        for n in np.flatnonzero(node_int[:, NI_HEIGHT] == 0):
            last[int(n)] = -1
    for i, (kind, idx, _, _) in enumerate(items):
        if kind != WORK_EVAL:
            last[int(idx)] = i
    for i, (kind, idx, _, _) in enumerate(items):
        deps = ()
        if kind == WORK_MERGE:
            deps = (int(node_int[idx, 5]), int(node_int[idx, 6]))
        elif kind == WORK_EVAL:
            deps = (int(job_root[idx]),)
        if any(d not in last or last[d] >= i for d in deps):
            return False
    return True


class PersistentFFBP:
    """A device work list plus workspace for ``bpmi_ffbp_persistent``.

    Node tables and arena are the caller's (device arrays of a flattened
    plan or of many patches).  Leaf inputs come either from ``sources`` (a
    device array of per-node ``bpmi_node_source`` records) or from the single
    patch ``data``/``pos``/``weight``/``phase``.  ``jobs``: host array of
    ``bpmi_eval_job`` records (``multipatch.EVAL_JOB``) with their root nodes.
    """

    def __init__(self, node_int_host, node_int, node_geo, offsets, arena, k, tile_cols_log2, *,
                 sources=None, data=None, pos=None, weight=None, phase=None, samples=0,
                 range_start=0.0, data_dr=1.0, jobs=None, job_roots=(), leaf_interp="lagrange4",
                 merge_interp="kaiser8", priority=None, max_wait_s=10.0, grid=0, leaves=True,
                 tiles=None):
        if not leaves:
            raise NotImplementedError(
                "a persistent work list without leaves is switched off: its one GPU run did not "
                "finish and the cause has not been found")
        node_int_host = np.asarray(node_int_host)
        self.nodes = len(node_int_host)
        self.njobs = 0 if jobs is None else len(jobs)
        self.leaves = bool(leaves)
        grids = [] if jobs is None else [(int(j["nx"]), int(j["ny"])) for j in jobs]
        items, node_tiles, job_blocks = work_items(node_int_host, tile_cols_log2,
                                                   [(r, nx, ny) for r, (nx, ny) in
                                                    zip(job_roots, grids)], priority,
                                                   leaves=self.leaves, tiles=tiles)
        if not check_order(items, node_int_host, job_roots, leaves=self.leaves):
            raise RuntimeError("persistent work list is not topologically ordered")
        self.items_host = items
        self.leaf_interp = _lib.interp_code(leaf_interp) if self.leaves else INTERP_NONE
        self.merge_interp = _lib.interp_code(merge_interp)
        self.grid = int(grid)
        keep = []

        def dev(a, dtype):
            d = DeviceArray.from_host(np.ascontiguousarray(a, dtype))
            keep.append(d)
            return d.ptr
        self.node_tiles_host = node_tiles

        nbytes = ctypes.c_size_t()
        _lib.call("bpmi_persistent_workspace_bytes", self.nodes, self.njobs, ctypes.byref(nbytes))
        self.workspace = DeviceArray.empty(-(-nbytes.value // 8))
        keep.append(self.workspace)
        sc = _lib.PersistentSchedule()
        sc.node_int, sc.node_geo, sc.offsets, sc.arena = node_int, node_geo, offsets, arena
        sc.sources = sources
        sc.data, sc.pos, sc.weight, sc.phase = data, pos, weight, phase
        sc.samples = int(samples)
        sc.tile_cols_log2 = int(tile_cols_log2)
        sc.range_start, sc.data_dr, sc.k = float(range_start), float(data_dr), float(k)
        if self.njobs:
            sc.jobs = dev(np.asarray(jobs).view(np.uint8), np.uint8)
            sc.job_root = dev(np.asarray(job_roots, np.int32), np.int32)
            sc.job_blocks = dev(job_blocks, np.int32)
        sc.njobs = self.njobs
        sc.items = dev(items, np.int32)
        sc.nitems = len(items)
        self.node_tiles_dev = DeviceArray.from_host(node_tiles)
        keep.append(self.node_tiles_dev)
        sc.node_tiles = self.node_tiles_dev.ptr
        sc.nodes = self.nodes
        sc.workspace = self.workspace.ptr
        sc.max_wait_s = float(max_wait_s)
        self.schedule = sc
        self._keep = keep

    @property
    def nitems(self):
        return int(self.schedule.nitems)

    def launch(self, stream=None):
        """Enqueue the workspace reset and the one launch (no synchronisation)."""
        _lib.call("bpmi_ffbp_persistent", ctypes.byref(self.schedule), self.leaf_interp,
                  self.merge_interp, self.grid, stream if stream is not None else stream_handle())

    def aborted(self):
        """After the launch completed: True when a dependency wait timed out."""
        v = ctypes.c_int()
        _lib.call("bpmi_persistent_aborted", _lib.c_void_p(self.workspace.ptr), self.nodes,
                  self.njobs, ctypes.byref(v))
        return bool(v.value)

    def completion_ms(self):
        """After the launch completed: each job's completion time after the
        first workgroup started, in ms (device wall clock)."""
        stamps = np.zeros(1 + self.njobs, np.uint64)
        _lib.call("bpmi_persistent_stamps", _lib.c_void_p(self.workspace.ptr), self.nodes,
                  self.njobs, stamps.ctypes.data_as(_lib.c_void_p))
        khz = ctypes.c_int()
        _lib.call("bpmi_wall_clock_khz", ctypes.byref(khz))
        return (stamps[1:].astype(np.int64) - np.int64(stamps[0])) / float(khz.value)

    def close(self):
        self._keep = []
        self.workspace = None
