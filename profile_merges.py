"""One C2-sized FFBP for hardware-counter collection.

    python profile_merges.py [--pulses 8192] [--no-untiled] [--cull grid]
                             [--merge-shapes W:C[,W:C...]] [--persistent]

Runs the split-precision FFBP of the 8192-pulse x 8192-bin stripmap workload
(leaf 16, factor 2, 8-tap Kaiser-sinc merges) once through the level-batched
program (LDS-tiled merges; ``--cull`` none, crop or grid selects the tiles as
the C2 step does, default grid: those the 4096 x 4096 Cartesian output needs),
then every merge level again through the
untiled merge kernel on the same node images, with ``--persistent`` the same
FFBP once more as one persistent launch (same culling), then one 1 GiB copy kernel
(`bpmi_crop` of a 16384 x 8192 complex64 image: 1 GiB read, 1 GiB written)
whose counters calibrate FETCH_SIZE / WRITE_SIZE.  Meant
to run under `rocprofv3 --pmc ...`; prints one line per phase.
"""
from __future__ import annotations

import argparse
import ctypes
import sys
import time
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent))

from bp_project_amd import _lib  # noqa: E402
from bp_project_amd.batched_resident_backend import (BatchedResidentProgram,  # noqa: E402
                                                     merge_block_table)
from bp_project_amd.device import DeviceArray, synchronize  # noqa: E402
from bp_project_amd.gpu_backend import HipBaseline, build_ffbp_plan  # noqa: E402
from bp_project_amd.scenes import polar_extent_for_grid, workload_scene  # noqa: E402

_vp = ctypes.c_void_p
COPY_BYTES = 1 << 30


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--pulses", type=int, default=8192)
    ap.add_argument("--no-untiled", action="store_true")
    ap.add_argument("--cull", default="grid", choices=["none", "crop", "grid"])
    ap.add_argument("--merge-shapes", default=None,
                    help="W:C (every level) or one W:C per level, comma-separated: the "
                         "merge launch shapes (occupancy hint, window cap)")
    ap.add_argument("--persistent", action="store_true",
                    help="also run the FFBP as one persistent launch (bpmi_ffbp_persistent)")
    a = ap.parse_args(argv)
    shapes = None
    if a.merge_shapes:
        shapes = [tuple(int(v) for v in item.split(":")) for item in a.merge_shapes.split(",")]
        if len(shapes) == 1:
            shapes = shapes[0]
    t = time.perf_counter()
    s = workload_scene("c2_ffbp_full", pulses=a.pulses, bins=a.pulses)
    extent = polar_extent_for_grid(s.grid, s.pos.mean(axis=0), margin=2.0)
    plan = build_ffbp_plan(s.pos, s.fc, extent, side=1, leaf_size=16, range_spacing=0.25,
                           angular_oversample=2.5)
    backend = HipBaseline(interp="lagrange4", merge_interp="kaiser8")
    dev = backend.upload(s.data, s.pos, s.weight, s.phase, plan)
    cull = {"none": None, "crop": "crop", "grid": s.grid}[a.cull]
    prog = BatchedResidentProgram(backend, plan, dev, "split", s.r_start, s.data_dr, s.k,
                                  cull=cull, merge_shapes=shapes)
    print(f"merge shapes: {prog.merge_shapes}", flush=True)
    if prog.cull:
        print(f"cull: {prog.cull}", flush=True)
    print(f"setup {time.perf_counter() - t:.1f} s", flush=True)
    prog.execute()
    synchronize()
    print("tiled FFBP done", flush=True)
    if not a.no_untiled:
        flat = prog.flat
        for height, nodes in enumerate(flat.levels):
            if height == 0:
                continue
            table = DeviceArray.from_host(
                np.asarray(merge_block_table(flat.node_int, nodes, "fp64"), np.int32))
            _lib.call("bpmi_ffbp_merges", _vp(prog._node_table.ptr),
                      _vp(prog._geometry_table.ptr), _vp(prog._offset_table.ptr),
                      _vp(prog.arena.ptr), _vp(table.ptr), table.shape[0], s.k,
                      _lib.MODE_SPLIT, _lib.interp_code("kaiser8"), None)
            synchronize()
        print("untiled merges done", flush=True)
    if a.persistent:
        pers = BatchedResidentProgram(backend, plan, dev, "split", s.r_start, s.data_dr, s.k,
                                      cull=cull, persistent=True)
        pers.execute()
        synchronize()
        same = np.array_equal(pers.output.get(), prog.output.get())
        print(f"persistent FFBP done ({pers.persistent.nitems} items), "
              f"output bitwise equal to the level launches: {same}", flush=True)
        pers.close()
    rows, cols = COPY_BYTES // (8 * 8192), 8192
    src, dst = DeviceArray.empty((rows, cols)), DeviceArray.empty((rows, cols))
    src.fill_zero()
    _lib.call("bpmi_crop", _vp(src.ptr), cols, 0, 0, rows, cols, _vp(dst.ptr), None)
    synchronize()
    print(f"calibration copy of {COPY_BYTES} bytes done", flush=True)
    prog.close()


if __name__ == "__main__":
    main()
