"""Throughput benchmark of the SAR FFBP / direct-BP hot path on one MI355X.

Default workload (``c2_ffbp_full``): FFBP of an 8192-pulse x 8192-bin
X-band stripmap scene (200 point targets + 20 dB clutter, seeded synthetic
data), leaf sub-apertures of 16 pulses merged by factor 2 over 9 levels with
8-tap Kaiser(beta=6)-sinc polar regridding, then evaluation onto a
4096 x 4096 Cartesian grid at 0.25 m.  One step = leaves + all merge levels +
Cartesian evaluation (reading the root's crop window in place), inputs
already resident in HBM.

value = effective Gpix*pulse/s = (output pixels x pulses) / step time, the
reference's direct-equivalent throughput (fast_bp/gpu_backend.py:205).

The JSON line also carries:
* ``quality``: the split-precision image against the FP64 image of the same
  algorithm computed on the GPU (NRMSE, peak-phase error at every target,
  worst PSLR/ISLR delta), plus the FFBP-vs-direct-BP error on sampled pixels;
* ``roofline``: achieved algorithmic FLOP/s of the dominant kernel (HIP events
  around its launches on the launch stream) against the MI355X FP32 vector
  peak -- backprojection is VALU-bound, it runs no MFMA;
* ``cpu_baseline``: the FP64 C/OpenMP restatement of the reference's Numba
  FFBP (the oracle) timed on the host cores on a bounded sample.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Gpix·pulse/s (BP, FFBP) at NRMSE≤1e-4 vs FP64; ms per autofocus grad step"
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of the dominant kernel, measured with rocprofv3
# FETCH_SIZE and WRITE_SIZE (separate runs), each scaled by its reading for a
# 1 GiB copy kernel (FETCH x 2.0, WRITE x 1.0); see
# profiles/ffbp_c2_merge_counters.md.  Keyed by (workload, mode, stage).
# The culled figure is from the merge with window tables; the unculled one
# from the same kernel before them (tables add 64 bytes per 32 x 32 tile).
MEASURED_TRAFFIC_BYTES = {  # (workload, mode, kernel, cull)
    ("c2_ffbp_full", "split", "batched_merge_tiled_kernel<kaiser8>", "none"): 4.22e9,
    ("c2_ffbp_full", "split", "batched_merge_tiled_kernel<kaiser8>", "grid"): 3.66e9,
}
# Schedule stage -> the HIP kernel its launches run (split mode, C2 settings).
STAGE_KERNEL = {
    "batched_bp_polar_split": "bp_split_kernel<lagrange4> (FFBP leaves)",
    "batched_merge_regrid_split": "batched_merge_tiled_kernel<kaiser8>",
    "eval_grid": "eval_grid_tiled_kernel<kaiser8>",
    "ffbp_persistent_split": "ffbp_persistent_kernel<lagrange4, kaiser8>",
}
# Algorithmic operation counts (add/sub/mul = 1, FMA = 2, sqrt/sin/cos = 1).
OPS_LEAF_LAGRANGE4 = 57      # per leaf pixel*pulse (4-point Lagrange)
OPS_DIRECT_LINEAR = 35       # per pixel*pulse (linear interpolation)
OPS_MERGE_KAISER8 = 742      # per parent pixel (two 8x8 Kaiser-sinc child regrids)
OPS_ADJOINT = 69             # per pixel*pulse (split adjoint, fourth-order fast path)
OPS_MERGE_LAGRANGE4 = 290    # per parent pixel (two 4x4 Lagrange child regrids)
OPS_EVAL_KAISER8 = 386       # per output pixel (one 8x8 regrid + remodulation)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=15)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--workload", default="c2_ffbp_full",
                   choices=["c2_ffbp_full", "c1_direct_small", "c3_direct_precision",
                            "c4_autofocus", "c5_multipatch", "range_compress"])
    p.add_argument("--mode", default="split")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-quality", action="store_true", help="skip the FP64 quality gate")
    p.add_argument("--leaf", type=int, default=16)
    p.add_argument("--range-spacing", type=float, default=0.25)
    p.add_argument("--angular-oversample", type=float, default=2.5)
    p.add_argument("--merge-interp", default="kaiser8")
    p.add_argument("--precision-study", action="store_true",
                   help="direct workloads: compare fp32/mixed/dfloat/split against fp64")
    p.add_argument("--cull", default="grid", choices=["none", "crop", "grid"],
                   help="C2: launch only the tiles the root crop ('crop') or the Cartesian "
                        "grid's footprint in it ('grid') depends on (output bitwise unchanged)")
    p.add_argument("--no-unculled", action="store_true",
                   help="C2 with culling: skip timing the unculled step beside it")
    p.add_argument("--rounds", type=int, default=15,
                   help="rounds of the randomized interleaved timing protocol")
    p.add_argument("--no-others", action="store_true",
                   help="C2 default run: skip the summaries of the other workloads")
    p.add_argument("--no-autotune", action="store_true",
                   help="C2: keep the default launch shapes instead of choosing the leaf pulse "
                        "block depth and one merge shape per level with the autotuner before "
                        "timing (C5 always tunes)")
    p.add_argument("--persistent", action="store_true",
                   help="FFBP workloads: one persistent launch for leaves and all merge levels "
                        "(C5: and the evaluation)")
    p.add_argument("--latency-groups", default="0.25,0.5,0.75",
                   help="C5: also serve the batch in latency groups cut at these fractions "
                        "of the patches ordered by work ('none': skip)")
    return p.parse_args()


class StageTimer:
    """HIP events between consecutive launches on the current stream."""

    def __init__(self):
        from bp_project_amd.device import Event
        self.Event = Event
        self.marks = []

    def mark(self, name):
        e = self.Event()
        e.record()
        self.marks.append((name, e))

    def durations(self):
        from bp_project_amd.device import elapsed_ms
        out = {}
        for (_, a), (name, b) in zip(self.marks[:-1], self.marks[1:]):
            out.setdefault(name, []).append(elapsed_ms(a, b))
        return out

    def step_durations(self, steps):
        """Per-step times: the marks after "start" split into ``steps`` equal runs."""
        from bp_project_amd.device import elapsed_ms
        seq = [elapsed_ms(a, b) for (_, a), (_, b) in zip(self.marks[:-1], self.marks[1:])]
        n = len(seq) // steps
        return [float(sum(seq[i * n:(i + 1) * n])) for i in range(steps)]


def step_timing(step_ms):
    """Median and interquartile range of per-step HIP-event times (consecutive
    warmed steps on the launch stream), the reference's reporting protocol
    (resident_experiment.py:87-119)."""
    v = np.asarray(step_ms, np.float64)
    q1, med, q3 = np.percentile(v, [25, 50, 75])
    return {"source": "HIP events around each timed step", "n": int(v.size),
            "median_ms": round(float(med), 4), "iqr_ms": round(float(q3 - q1), 4),
            "min_ms": round(float(v.min()), 4), "max_ms": round(float(v.max()), 4)}


def interleaved_timing(variants, rounds=15, warmups=2, seed=20260917):
    """The reference's timing protocol (fast_bp/resident_experiment.py:87-119):
    ``warmups`` untimed calls of every variant, then ``rounds`` rounds that
    each call every variant once in a seeded random order; each call is
    bracketed by HIP events on the launch stream and by the host wall clock,
    and synchronised.  Returns per-variant median/IQR of both and the orders."""
    from bp_project_amd.device import Event, elapsed_ms
    rng = np.random.default_rng(seed)
    names = list(variants)
    for name in names:
        for _ in range(warmups):
            variants[name]()
    ev = {n: [] for n in names}
    wall = {n: [] for n in names}
    orders = []
    a, b = Event(), Event()
    for _ in range(rounds):
        order = [names[i] for i in rng.permutation(len(names))]
        orders.append(order)
        for name in order:
            t = time.perf_counter()
            a.record()
            variants[name]()
            b.record()
            b.synchronize()
            wall[name].append(1000 * (time.perf_counter() - t))
            ev[name].append(elapsed_ms(a, b))

    def stats(v):
        q1, med, q3 = np.percentile(np.asarray(v), [25, 50, 75])
        return round(float(med), 4), round(float(q3 - q1), 4)
    out = {"protocol": "seeded randomized interleaved rounds (resident_experiment.py:87-119)",
           "rounds": rounds, "warmups": warmups, "seed": seed, "orders": orders}
    for n in names:
        (em, ei), (wm, wi) = stats(ev[n]), stats(wall[n])
        out[n] = {"event_median_ms": em, "event_iqr_ms": ei, "wall_median_ms": wm,
                  "wall_iqr_ms": wi}
    return out


_PEAKS = None


def measured_peaks(blocks=2048, iters=2048, copy_bytes=1 << 30, repeats=5):
    """The peaks the roofline divides by, re-measured on this GPU (HIP events,
    median of ``repeats`` after a warm-up): the FP32 FMA issue rate of
    ``fma_peak_kernel`` (vector FMAs, packed FP32 not used, as in every
    kernel of this build) and a device copy of ``copy_bytes`` (``bpmi_crop``
    of a complex64 image: that many bytes read and written).  Cached per
    process."""
    global _PEAKS
    if _PEAKS is not None:
        return _PEAKS
    from bp_project_amd import _lib
    from bp_project_amd.device import DeviceArray, Event, elapsed_ms, stream_handle, synchronize
    vp = _lib.c_void_p
    out = DeviceArray.empty((blocks * 256,), np.float32)
    rows, cols = copy_bytes // (8 * 8192), 8192
    src, dst = DeviceArray.empty((rows, cols)), DeviceArray.empty((rows, cols))
    src.fill_zero()

    def fma():
        _lib.call("bpmi_probe_fma_peak", vp(out.ptr), blocks, iters, stream_handle())

    def copy():
        _lib.call("bpmi_crop", vp(src.ptr), cols, 0, 0, rows, cols, vp(dst.ptr), stream_handle())

    def timed(fn):
        fn()
        synchronize()
        ms = []
        for _ in range(repeats):
            a, b = Event(), Event()
            a.record()
            fn()
            b.record()
            b.synchronize()
            ms.append(elapsed_ms(a, b))
        return float(np.median(ms))
    t_fma, t_copy = timed(fma), timed(copy)
    flops = blocks * 256 * iters * 16 * 8 * 2
    _PEAKS = {"fp32_fma_tflops": round(flops / (t_fma * 1e-3) / 1e12, 2),
              "copy_gbs": round(2 * rows * cols * 8 / (t_copy * 1e-3) / 1e9, 1),
              "method": f"fma_peak_kernel ({blocks} x 256 threads, 8 independent v_fma_f32 "
                        f"chains, {flops:.3g} FLOPs) and a {copy_bytes >> 20} MiB device copy "
                        "(read + write), HIP events, median of 5"}
    del out, src, dst
    return _PEAKS


def attach_measured(roof):
    """Add the re-measured peaks to a roofline record and the achieved rate's
    fraction of them (FP32 FMA rate for TFLOP/s records, copy bandwidth for
    GB/s records)."""
    if not roof:
        return roof
    pk = measured_peaks()
    roof.setdefault("measured_peak", pk)
    if roof.get("unit") == "TFLOP/s":
        roof["frac_of_measured_fp32"] = round(roof["achieved"] / pk["fp32_fma_tflops"], 4)
    elif roof.get("unit") == "GB/s":
        roof["frac_of_measured_copy"] = round(roof["achieved"] / pk["copy_gbs"], 4)
    return roof


def step_attribution(stages, steps, parts):
    """The FFBP step's roofline attributed per stage: for each stage the time
    its algorithmic ops need at the FP32 vector peak and its bytes need at
    the HBM peak, the larger of the two, summed over the stages, against the
    measured step.  ``parts``: {stage: (ops per step, bytes per step)}."""
    out, ideal = {}, 0.0
    for name, (ops, nbytes) in parts.items():
        if name not in stages:
            continue
        t_ops = ops / (PEAK_FP32_TFLOPS * 1e12) * 1e3
        t_bytes = nbytes / (PEAK_HBM_GBS * 1e9) * 1e3
        ms = stages[name]["ms_per_step"]
        out[name] = {"ops": float(ops), "bytes": float(nbytes), "bound_ms": round(max(t_ops, t_bytes), 4),
                     "bound": "valu" if t_ops >= t_bytes else "hbm", "measured_ms": round(ms, 4),
                     "frac": round(max(t_ops, t_bytes) / ms, 4)}
        ideal += max(t_ops, t_bytes)
    measured = sum(stages[n]["ms_per_step"] for n in out)
    return {"stages": out, "bound_ms": round(ideal, 4), "measured_ms": round(measured, 4),
            "frac": round(ideal / measured, 4) if measured else None,
            "method": "sum over stages of max(ops / FP32 peak, bytes / HBM peak) over the "
                      "measured stage times (merges: measured HBM bytes; leaves, evaluation: "
                      "compulsory bytes)"}


def autotune_summary(config, report, cached, defaults, per_stage=("merge_shapes",)):
    """The tuner's choice, its per-candidate stage times (ms, medians of
    event-timed runs) and, per knob group, the timed stages' total under the
    default shape against the chosen one."""
    gain = {}
    for name, r in report.items():
        times = r.get("times") or {}
        d = json.dumps(defaults.get(name))
        choice = r.get("choice", config.get(name))
        if d not in times:
            continue
        if name in per_stage:
            tuned = sum(times[json.dumps(c)][i] for i, c in enumerate(choice))
        else:
            tuned = sum(times[json.dumps(choice)])
        gain[name] = {"default_ms": round(float(sum(times[d])), 4),
                      "tuned_ms": round(float(tuned), 4)}
    return {"config": config, "cached": bool(cached),
            "search": "coordinate search, one knob group at a time with the others at their "
                      "best; each launch stage timed by HIP events between launches",
            "candidates_ms": {name: {"stages": r.get("stages"),
                                     "times": {k: [round(float(x), 4) for x in v]
                                               for k, v in (r.get("times") or {}).items()}}
                              for name, r in report.items()},
            "default_vs_tuned": gain}


def ffbp_workload(args, scene):
    from bp_project_amd import _lib
    from bp_project_amd.batched_resident_backend import BatchedResidentProgram
    from bp_project_amd.device import DeviceArray, stream_handle
    from bp_project_amd.gpu_backend import HipBaseline, build_ffbp_plan, geo_array
    from bp_project_amd.scenes import polar_extent_for_grid

    s = scene
    g = s.grid
    extent = polar_extent_for_grid(g, s.pos.mean(axis=0), margin=2.0)
    t0 = time.perf_counter()
    plan = build_ffbp_plan(s.pos, s.fc, extent, side=1, leaf_size=args.leaf,
                           range_spacing=args.range_spacing,
                           angular_oversample=args.angular_oversample)
    log(f"plan: {plan.counts['nodes']} nodes, root {plan.root.nr}x{plan.root.nu}, crop "
        f"{plan.shape}, {time.perf_counter() - t0:.2f} s")
    backend = HipBaseline(interp="lagrange4", merge_interp=args.merge_interp)
    dev = backend.upload(s.data, s.pos, s.weight, s.phase, plan)
    vp = _lib.c_void_p

    def make(mode, cull_region=None):
        cull = None
        if mode == "split" and cull_region not in (None, "none"):
            cull = "crop" if cull_region == "crop" else g
        prog = BatchedResidentProgram(backend, plan, dev, mode, s.r_start, s.data_dr, s.k,
                                      persistent=args.persistent and mode == "split", cull=cull)
        cart = DeviceArray.empty(g.shape)
        geo = geo_array(plan)
        interp = _lib.interp_code(args.merge_interp)
        code = _lib.mode_code(mode)
        # The Cartesian evaluation reads the cropped root in place: the crop
        # is a strided view (row length = root nu) instead of a copy.
        root = prog.node_image(plan.root.index)
        crop_ptr = root.ptr + 8 * (plan.crop_r[0] * plan.root.nu + plan.crop_u[0])

        def evaluate():
            _lib.call("bpmi_eval_grid", vp(crop_ptr), geo.ctypes.data_as(vp), plan.nr, plan.nu,
                      plan.root.nu, g.x0, g.dx, g.nx, g.y0, g.dy, g.ny, s.k, code, interp,
                      vp(cart.ptr), stream_handle())

        def step(timer=None):
            # the schedule is read at each call: set_launch_shapes() edits it
            for name, launcher, grid_, block, a in prog._schedule:
                if name == "crop_float2":
                    continue
                launcher(grid_, block, a)
                if timer:
                    timer.mark(name)
            evaluate()
            if timer:
                timer.mark("eval_grid")
        return prog, cart, step, geo

    return plan, backend, dev, make


def run_ffbp(args, scene):
    from bp_project_amd.device import Event, elapsed_ms, synchronize
    from bp_project_amd.metrics import fidelity, peak_phase_error, pslr_islr_deltas, target_peaks

    plan, backend, dev, make = ffbp_workload(args, scene)
    g = scene.grid
    prog, cart, step, _ = make(args.mode, args.cull)
    tuned = None
    if not args.no_autotune and args.mode == "split" and not args.persistent:
        from types import SimpleNamespace
        t = time.perf_counter()
        cfg, report, cached = prog.autotune(
            execute=lambda mark: step(None if mark is None else SimpleNamespace(mark=mark)))
        tuned = autotune_summary(cfg, report, cached,
                                 {"pulse_block": 32, "merge_shapes": [4, 2176]})
        log(f"autotuned in {time.perf_counter() - t:.1f} s:", json.dumps(tuned["config"]),
            json.dumps(tuned["default_vs_tuned"]))
    for _ in range(args.warmup):
        step()
    synchronize()
    timer = StageTimer()
    start, end = Event(), Event()
    synchronize()
    start.record()
    timer.mark("start")
    for _ in range(args.steps):
        step(timer)
    end.record()
    synchronize()
    ms_total = elapsed_ms(start, end)
    ms_step = ms_total / args.steps
    stages = timer.durations()
    step_ms = timer.step_durations(args.steps)
    # per-stage algorithmic work
    flat = prog.flat
    nodes_px = flat.node_int[:, 2].astype(np.int64) * flat.node_int[:, 3]
    if prog.computed_pixels is not None:
        nodes_px = np.array([prog.computed_pixels[i] for i in range(flat.node_count)], np.int64)
    spans = (flat.node_int[:, 1] - flat.node_int[:, 0]).astype(np.int64)
    leaf_pp = int(sum(nodes_px[i] * spans[i] for i in flat.levels[0]))
    merge_px = int(sum(nodes_px[i] for lvl in flat.levels[1:] for i in lvl))
    ops_merge = OPS_MERGE_KAISER8 if args.merge_interp == "kaiser8" else OPS_MERGE_LAGRANGE4
    work = {
        f"batched_bp_polar_{args.mode}": leaf_pp * OPS_LEAF_LAGRANGE4,
        f"batched_merge_regrid_{args.mode}": merge_px * ops_merge,
        f"ffbp_persistent_{args.mode}": merge_px * ops_merge + leaf_pp * OPS_LEAF_LAGRANGE4,
        "eval_grid": g.nx * g.ny * OPS_EVAL_KAISER8,
    }
    per_stage = {}
    for name, vals in stages.items():
        per_stage[name] = {"launches_per_step": len(vals) // args.steps,
                           "ms_per_step": float(np.sum(vals) / args.steps),
                           "mean_launch_ms": float(np.mean(vals))}
    dominant = max((n for n in per_stage if n in work and work[n] > 0),
                   key=lambda n: per_stage[n]["ms_per_step"])
    dom_ms = per_stage[dominant]["ms_per_step"]
    achieved = work[dominant] / (dom_ms * 1e-3) / 1e12
    log("stages:", json.dumps(per_stage))
    result = {
        "ms_per_step": ms_step,
        "value": g.nx * g.ny * scene.pulses / (ms_step * 1e-3) / 1e9,
        "stages": per_stage,
        "timing": step_timing(step_ms),
        "roofline": {
            "bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
            "traffic": MEASURED_TRAFFIC_BYTES.get((args.workload, args.mode,
                                                   STAGE_KERNEL.get(dominant), args.cull)),
            "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                            "copy-kernel calibrated; profiles/ffbp_c2_merge_counters.md)",
            "kernel": STAGE_KERNEL.get(dominant, dominant),
            "stage": dominant,
            "ops_per_unit": (OPS_LEAF_LAGRANGE4 if "bp_polar" in dominant else
                             ops_merge if "merge" in dominant else
                             f"{OPS_LEAF_LAGRANGE4} per leaf pixel*pulse + {ops_merge} per parent "
                             "pixel" if "persistent" in dominant else OPS_EVAL_KAISER8),
            "unit_of_work": ("leaf pixel*pulse" if "bp_polar" in dominant else
                             "parent pixel" if "merge" in dominant else
                             "leaves + merges" if "persistent" in dominant else "output pixel"),
            "launches_per_step": per_stage[dominant]["launches_per_step"],
            "note": "FP32 vector-ALU bound (no MFMA); peak = MI355X FP32 vector 157.3 "
                    "TFLOP/s (packed); packed FP32 is disabled in this build, so FMA "
                    "code tops out at the unpacked 78.6 TFLOP/s",
        },
        "config": {
            "workload": "c2_ffbp_full", "pulses": scene.pulses, "bins": scene.bins,
            "targets": len(scene.targets), "leaf_size": args.leaf, "merge_factor": 2,
            "levels": len(flat.levels) - 1, "merge_interp": args.merge_interp,
            "leaf_interp": "lagrange4", "range_spacing_m": args.range_spacing,
            "angular_oversample": args.angular_oversample, "output": [g.ny, g.nx],
            "pixel_m": g.dx, "mode": args.mode, "root_polar": [plan.root.nr, plan.root.nu],
            "persistent": bool(args.persistent),
            "arena_gb": round(flat.arena_elems * 8 / 1e9, 2),
            "cull": args.cull,
        },
    }
    attach_measured(result["roofline"])
    leaf_px = int(sum(nodes_px[i] for i in flat.levels[0]))
    merge_traffic = MEASURED_TRAFFIC_BYTES.get((args.workload, args.mode,
                                                STAGE_KERNEL.get(f"batched_merge_regrid_{args.mode}"),
                                                args.cull))
    result["roofline"]["step"] = step_attribution(per_stage, args.steps, {
        f"batched_bp_polar_{args.mode}": (
            leaf_pp * OPS_LEAF_LAGRANGE4, 8 * leaf_px + 8 * scene.pulses * scene.bins),
        f"batched_merge_regrid_{args.mode}": (
            merge_px * ops_merge,
            (merge_traffic * (len(flat.levels) - 1)) if merge_traffic else 16 * merge_px),
        "eval_grid": (g.nx * g.ny * OPS_EVAL_KAISER8,
                      8 * g.nx * g.ny + 8 * int(nodes_px[plan.root.index])),
    })
    if tuned is not None:
        result["autotune"] = tuned
        if "merge" in dominant:
            # the merge launches run the autotuned <interp, waves, cap> instance per level
            result["roofline"]["kernel_instances"] = [
                f"batched_merge_tiled_kernel<{args.merge_interp}, {w}, {c}>"
                for w, c in prog.merge_shapes]
    if prog.cull is not None:
        result["cull"] = dict(prog.cull, computed_fraction=round(
            prog.cull["computed_pixels"] / prog.cull["node_pixels"], 4))
    if prog.cull is not None and not args.no_unculled:
        # the same step without culling, timed against the culled step by
        # the reference's randomized interleaved protocol; outputs compared
        # bit for bit
        full_prog, full_cart, full_step, _ = make(args.mode)
        if tuned is not None:
            full_prog.set_launch_shapes(prog.pulse_block, prog.merge_shapes)
        inter = interleaved_timing({"culled": step, "unculled": full_step},
                                   rounds=args.rounds, warmups=args.warmup)
        step()
        synchronize()
        same = bool(np.array_equal(cart.get(), full_cart.get()))
        full_prog.close()
        result["timing"]["interleaved"] = inter
        result["cull"].update(unculled_ms_per_step=inter["unculled"]["event_median_ms"],
                              output_bitwise_equal_unculled=same)
        log("interleaved:", json.dumps({k: v for k, v in inter.items() if k != "orders"}))
        if not same:
            raise SystemExit("culled output differs from the unculled output")
    if not args.no_quality:
        img = cart.get()
        ref_prog, ref_cart, ref_step, _ = make("fp64")
        ref_step()
        synchronize()
        ref = ref_cart.get()
        ref_prog.close()
        f = fidelity(ref, img)
        peaks = target_peaks(np.abs(ref), g, scene.targets)
        ph = peak_phase_error(ref, img, peaks)
        pis = pslr_islr_deltas(ref, img, g, peaks[:12])
        # FFBP vs direct BP (algorithm error) on 2048 sampled pixels
        from bp_project_amd.ffbp import backproject_points
        rng = np.random.default_rng(7)
        sel = rng.choice(g.nx * g.ny, 2048, replace=False)
        sel = np.unique(np.r_[sel, [p[0] * g.nx + p[1] for p in peaks]])
        pts = g.points()[sel]
        direct = backproject_points(scene.data, scene.pos, pts, scene.r_start, scene.data_dr,
                                    scene.k, scene.weight, scene.phase, mode="fp64")
        alg = float(np.linalg.norm(ref.ravel()[sel] - direct) / np.linalg.norm(direct))
        result["quality"] = {
            "reference": "FP64 FFBP of the same plan on the GPU (fp64 mode)",
            "nrmse": f["relative_complex_error"], "ser_db": f["ser_db"],
            "max_peak_phase_error_rad": float(max(ph)) if ph else None,
            "targets_measured": len(ph),
            "max_abs_pslr_islr_delta_db": pis["max_abs_delta_db"],
            "ffbp_vs_direct_fp64_error_sampled": alg,
            "passes_nrmse_1e-4": bool(f["relative_complex_error"] <= 1e-4),
        }
        log("quality:", json.dumps(result["quality"]))
        if not result["quality"]["passes_nrmse_1e-4"]:
            raise SystemExit(f"quality gate failed: NRMSE {f['relative_complex_error']:.3e} "
                             "> 1e-4 against the FP64 image; no throughput is reported")
    prog.close()
    return result


def cpu_baseline_ffbp(args, scene):
    """Oracle (FP64 C/OpenMP FFBP + evaluation) on a 1024 x 1024 sub-patch with
    all pulses: one untimed run, then the median of three timed runs."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    from bp_project_amd.scenes import CartesianGrid, polar_extent_for_grid
    g = scene.grid
    n = 1024
    cx = g.x0 + g.dx * g.nx / 2
    cy = g.y0 + g.dy * g.ny / 2
    sub = CartesianGrid(x0=cx - n * g.dx / 2, dx=g.dx, nx=n, y0=cy - n * g.dy / 2, dy=g.dy, ny=n)
    extent = polar_extent_for_grid(sub, scene.pos.mean(axis=0), margin=2.0)
    pts = sub.points()

    def run():
        t = time.perf_counter()
        res, _ = oracle.form_ffbp(scene.data, scene.pos, scene.r_start, scene.data_dr, scene.fc,
                                  extent, side=1, leaf_size=args.leaf,
                                  range_spacing=args.range_spacing,
                                  angular_oversample=args.angular_oversample,
                                  weight=scene.weight, phase=scene.phase, policy="parity",
                                  leaf_interp="lagrange4", merge_interp=args.merge_interp)
        oracle.evaluate_image(res.image, res.origin, res.r0, res.dr, res.u0, res.du, pts,
                              scene.k, interp=args.merge_interp)
        return time.perf_counter() - t

    run()
    times = [run() for _ in range(3)]
    dt = float(np.median(times))
    rate = n * n * scene.pulses / dt / 1e9
    return {"value": round(rate, 4), "unit": "Gpix·pulse/s", "cores": oracle.num_threads(),
            "kind": "port", "seconds": round(dt, 2), "seconds_runs": [round(x, 2) for x in times],
            "sample": f"FP64 oracle FFBP (same leaf/merge parameters, {args.merge_interp} merges) "
                      f"+ evaluate_image of a {n}x{n} sub-patch at {g.dx} m with all "
                      f"{scene.pulses} pulses, one untimed run then the median of 3; "
                      f"effective px*pulse/s"}


def run_direct(args, scene, name):
    from bp_project_amd import _lib
    from bp_project_amd.device import DeviceArray, Event, elapsed_ms, synchronize
    from bp_project_amd.metrics import fidelity, peak_phase_error, target_peaks
    vp = _lib.c_void_p
    g = scene.grid
    interp = "linear" if name == "c1_direct_small" else "lagrange4"
    d = DeviceArray.from_host(scene.data)
    p = DeviceArray.from_host(scene.pos)
    out = DeviceArray.empty(g.shape)
    # small grids: pulse-split launches through a workspace of partial images
    ws_bytes = ctypes.c_size_t()
    _lib.call("bpmi_bp_grid_workspace_bytes", g.nx, g.ny, scene.pulses, ctypes.byref(ws_bytes))
    ws = DeviceArray.empty((ws_bytes.value // 8,)) if ws_bytes.value else None

    def step(mode=args.mode, dst=out):
        _lib.call("bpmi_bp_grid_ws", vp(d.ptr), scene.bins, vp(p.ptr), None, None, 0,
                  scene.pulses, g.x0, g.dx, g.nx, g.y0, g.dy, g.ny, 0.0, scene.r_start,
                  scene.data_dr, scene.k, _lib.mode_code(mode), _lib.interp_code(interp), 0,
                  vp(dst.ptr), vp(ws.ptr) if ws else None, ws_bytes.value, None)
    for _ in range(args.warmup):
        step()
    synchronize()
    start, end = Event(), Event()
    start.record()
    for _ in range(args.steps):
        step()
    end.record()
    synchronize()
    ms = elapsed_ms(start, end) / args.steps
    pp = g.nx * g.ny * scene.pulses
    ops = OPS_DIRECT_LINEAR if interp == "linear" else OPS_LEAF_LAGRANGE4
    achieved = pp * ops / (ms * 1e-3) / 1e12
    result = {"ms_per_step": ms, "value": pp / (ms * 1e-3) / 1e9,
              "roofline": {"bound": "valu", "achieved": round(achieved, 3),
                           "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                           "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
                           "kernel": "bp_split_kernel", "ops_per_unit": ops,
                           "unit_of_work": "pixel*pulse",
                           "launches_per_step": 2 if ws is not None else 1,
                           "note": "FP32 vector-ALU bound (no MFMA)"},
              "config": {"workload": name, "pulses": scene.pulses, "bins": scene.bins,
                         "output": [g.ny, g.nx], "pixel_m": g.dx, "interp": interp,
                         "mode": args.mode,
                         "pulse_split_partials": ws_bytes.value // (g.nx * g.ny * 8)}}
    if not args.no_quality or args.precision_study:
        from bp_project_amd.metrics import pslr_islr_deltas
        img = out.get()
        ref_out = DeviceArray.empty(g.shape)
        t0 = time.perf_counter()
        step("fp64", ref_out)
        synchronize()
        log(f"fp64 reference image: {time.perf_counter() - t0:.2f} s")
        ref = ref_out.get()
        peaks = target_peaks(np.abs(ref), g, scene.targets)

        def assess(image):
            f = fidelity(ref, image)
            ph = peak_phase_error(ref, image, peaks)
            pis = pslr_islr_deltas(ref, image, g, peaks[:16])
            return {"nrmse": f["relative_complex_error"],
                    "max_peak_phase_error_rad": float(max(ph)) if ph else None,
                    "max_abs_pslr_islr_delta_db": pis["max_abs_delta_db"],
                    "passes_nrmse_1e-4": bool(f["relative_complex_error"] <= 1e-4)}
        result["quality"] = dict(reference="FP64 direct BP on the GPU", **assess(img))
        log("quality:", json.dumps(result["quality"]))
        if args.mode != "fp32" and not result["quality"]["passes_nrmse_1e-4"]:
            raise SystemExit(f"quality gate failed: {args.mode} NRMSE "
                             f"{result['quality']['nrmse']:.3e} > 1e-4 against the FP64 image")
        if args.precision_study:
            study = {}
            for mode in ("fp32", "mixed", "dfloat", "split"):
                dst = DeviceArray.empty(g.shape)
                t0 = time.perf_counter()
                start, end = Event(), Event()
                start.record()
                step(mode, dst)
                end.record()
                synchronize()
                ms_mode = elapsed_ms(start, end)
                study[mode] = dict(ms=round(ms_mode, 3),
                                   gpix_pulse_per_s=round(pp / (ms_mode * 1e-3) / 1e9, 2),
                                   **assess(dst.get()))
                log(f"precision study {mode}: {json.dumps(study[mode])}")
            result["precision_study"] = study
    return result


def run_autofocus(args, scene):
    """Config 4: 30 Adam iterations of entropy autofocus through the HIP adjoint,
    against plain FP64 BP with finite-difference gradients: the FD cost per
    parameter measured on a block of pulses at full size, and FD autofocus run
    to the adjoint's final entropy on a reduced scene of the same geometry."""
    import torch
    from bp_project_amd.autofocus import GridOperator, entropy, gradient_autofocus
    from bp_project_amd.device import Event, elapsed_ms, synchronize
    op = GridOperator.from_scene(scene, mode=args.mode, interp="lagrange4")
    res = gradient_autofocus(op, scene.phase, scene.pos, iterations=30, lr_phase=0.05,
                             lr_pos=5e-4)
    steps_ms = res["ms_per_iteration"][2:] or res["ms_per_iteration"]
    ms_step = float(np.median(steps_ms))
    phase = torch.from_numpy(scene.phase).cuda()
    pos = torch.from_numpy(scene.pos).cuda()
    # the two HIP launches of a gradient step on their own: split forward BP, adjoint
    rng = np.random.default_rng(1)
    gimg = torch.from_numpy((rng.standard_normal((scene.grid.ny, scene.grid.nx)) + 1j *
                             rng.standard_normal((scene.grid.ny, scene.grid.nx)))
                            .astype(np.complex64)).cuda()
    a, b = Event(), Event()
    kernel_ms = {}
    for name, fn in (("forward_bp_ms", lambda: op.image(phase, pos)),
                     ("adjoint_ms", lambda: op.adjoint(phase, pos, gimg)),
                     ("fp64_bp_image_ms", lambda: op.image64(phase, pos))):
        fn()
        synchronize()
        a.record()
        for _ in range(5):
            fn()
        b.record()
        synchronize()
        kernel_ms[name] = round(elapsed_ms(a, b) / 5, 3)
    fd = fd_measured_block(op, scene, block=8)
    fd_step_ms = fd["ms_per_parameter"] * 4 * scene.pulses
    with torch.no_grad():
        true_img = op.image(torch.from_numpy(scene.phase - scene.meta["phase_error"]).cuda(),
                            torch.from_numpy(scene.true_pos).cuda())
        true_entropy = float(entropy(true_img))
    reduced = fd_equal_entropy(args)
    g = scene.grid
    pp = g.nx * g.ny * scene.pulses
    achieved = pp * OPS_ADJOINT / (kernel_ms["adjoint_ms"] * 1e-3) / 1e12
    return {
        "ms_per_step": ms_step,
        "value": pp / (ms_step * 1e-3) / 1e9,
        "autofocus": {
            "ms_per_gradient_step": round(ms_step, 3),
            "iterations": len(res["ms_per_iteration"]),
            "entropy_initial": res["entropy"][0], "entropy_final": res["best_entropy"],
            "entropy_with_true_phase": true_entropy, **kernel_ms,
            "fd_measured_block": fd,
            "fd_fp64_gradient_step_ms": round(fd_step_ms, 1),
            "fd_fp64_gradient_step_source": (
                f"measured: central differences over {fd['parameters']} parameters "
                f"({fd['pulses']} pulses x phase, x, y, z; {fd['images']} FP64 images) took "
                f"{fd['ms']:.1f} ms = {fd['ms_per_parameter']:.3f} ms per parameter; "
                f"x {4 * scene.pulses} parameters of a full gradient"),
            "speedup_vs_fd_fp64_per_step": round(fd_step_ms / ms_step, 1),
            "equal_entropy_reduced_scene": reduced},
        "roofline": {
            "bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
            "kernel": "bp_adjoint_split_kernel (+ adjoint_reduce_kernel)",
            "ops_per_unit": OPS_ADJOINT, "unit_of_work": "pixel*pulse",
            "note": "dominant kernel of a gradient step (the forward BP is the SPLIT "
                    "kernel of C3); time = HIP events around one adjoint call"},
        "config": {"workload": "c4_autofocus", "pulses": scene.pulses, "bins": scene.bins,
                   "output": [g.ny, g.nx], "pixel_m": g.dx, "mode": args.mode,
                   "optimizer": "Adam(phase lr 0.05, pos lr 5e-4), 30 iterations"},
    }


def fd_measured_block(op, scene, block=8):
    """Central-difference entropy gradient of ``block`` pulses (phase and
    x, y, z each: 8 FP64-mode images per pulse) at full size, timed."""
    from bp_project_amd.autofocus import finite_difference_gradient
    from bp_project_amd.device import synchronize
    pulses = np.random.default_rng(3).choice(scene.pulses, block, replace=False)
    finite_difference_gradient(op, scene.phase, scene.pos, pulses[:1], h=1e-4, mode="fp64")
    synchronize()
    t = time.perf_counter()
    finite_difference_gradient(op, scene.phase, scene.pos, pulses, h=1e-4, mode="fp64")
    synchronize()
    ms = 1000 * (time.perf_counter() - t)
    return {"pulses": block, "parameters": 4 * block, "images": 8 * block, "ms": round(ms, 2),
            "ms_per_parameter": round(ms / (4 * block), 4),
            "timing": "host wall clock around the whole FD evaluation (images, entropies)"}


def fd_equal_entropy(args, iterations=30):
    """A reduced scene of the C4 geometry (256 pulses onto 64 x 64 at 0.5 m,
    same phase random walk and along-track error): the adjoint autofocus runs
    ``iterations`` Adam steps; FD autofocus (fd_autofocus: same optimiser,
    FP64 images, 8 P images per step) runs until it reaches the adjoint's
    final entropy (at most twice as many steps).  Both final states are
    scored with the same FP64 image."""
    import torch
    from bp_project_amd.autofocus import (GridOperator, entropy, fd_autofocus,
                                          gradient_autofocus)
    from bp_project_amd.scenes import stripmap_scene
    s = stripmap_scene(pulses=256, bins=1024, n_targets=8, grid_n=64, grid_spacing=0.5,
                       seed=20260920, phase_walk_rms=0.5, along_track_error=0.02)
    op = GridOperator.from_scene(s, mode=args.mode, interp="lagrange4")
    kw = dict(lr_phase=0.05, lr_pos=5e-4)
    gradient_autofocus(op, s.phase, s.pos, iterations=2, **kw)  # warm-up
    adj = gradient_autofocus(op, s.phase, s.pos, iterations=iterations, **kw)
    target = adj["entropy"][-1]
    reach = next((i for i, e in enumerate(adj["entropy"]) if e <= target), iterations)
    adj_ms = float(np.sum(adj["ms_per_iteration"][:reach]))
    fd = fd_autofocus(op, s.phase, s.pos, iterations=2 * iterations, target=target, **kw)
    fd_ms = float(np.sum(fd["ms_per_iteration"]))
    fd_steps = len(fd["ms_per_iteration"])
    reached = fd["entropy"][-1] <= target

    def score(ph, ps):
        with torch.no_grad():
            return float(entropy(op.image64(torch.from_numpy(ph).cuda(),
                                            torch.from_numpy(ps).cuda())))
    return {"scene": "256 pulses x 1024 bins onto 64 x 64 at 0.5 m, 8 targets, phase walk "
                     "0.5 rad RMS, along-track error 2 cm (seed 20260920)",
            "target_entropy": target, "adjoint_steps_to_target": reach,
            "adjoint_ms_to_target": round(adj_ms, 2),
            "fd_steps": fd_steps, "fd_reached_target": bool(reached),
            "fd_ms": round(fd_ms, 1), "fd_images": fd["images"],
            "fd_ms_per_step": round(fd_ms / max(fd_steps, 1), 2),
            "adjoint_ms_per_step": round(float(np.median(adj["ms_per_iteration"])), 3),
            "final_entropy_fp64_image": {"adjoint": score(adj["phase"], adj["pos"]),
                                         "fd": score(fd["phase"], fd["pos"])},
            "speedup_to_equal_entropy": round(fd_ms / max(adj_ms, 1e-9), 1),
            "note": "measured wall time of both loops (synchronised per step); the FD run "
                    "stops at the first step whose entropy is at most the target"}


def cpu_baseline_autofocus(scene, budget_s=10.0):
    """The oracle's FP64 gradient step -- forward BP, entropy gradient
    image, adjoint over every pulse -- on a subset of grid rows (the image
    gradient of those rows; the work per row does not depend on its values)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    g = scene.grid

    def run(rows):
        pts = g.points()[: rows * g.nx]
        t = time.perf_counter()
        img = oracle.backproject_points(scene.data, scene.pos, pts, scene.r_start,
                                        scene.data_dr, scene.k, scene.weight, scene.phase)
        _, grad = oracle.entropy_and_grad(img)
        oracle.backproject_points_adjoint(scene.data, scene.pos, pts, scene.r_start,
                                          scene.data_dr, scene.k, scene.weight, scene.phase,
                                          grad)
        return len(pts), time.perf_counter() - t
    n, dt = run(1)
    rows = min(g.ny, max(1, int(budget_s / max(dt, 1e-3))))
    n, dt = run(rows)
    return {"value": round(n * scene.pulses / dt / 1e9, 4), "unit": "Gpix·pulse/s",
            "cores": oracle.num_threads(), "kind": "port", "seconds": round(dt, 2),
            "ms_per_gradient_step_extrapolated": round(1000 * dt * g.ny / rows, 1),
            "sample": f"FP64 oracle forward BP + entropy gradient + adjoint of {rows} grid rows "
                      f"({n} pixels) x {scene.pulses} pulses; rate in pixel*pulse/s of a "
                      "gradient step (extrapolation to the full grid by rows)"}


def run_multipatch(args):
    """Config 5: 64 uneven patches, one HIP graph per batch vs per-patch launches."""
    from bp_project_amd.device import Event, elapsed_ms, synchronize
    from bp_project_amd.multipatch import MultiPatchProgram, make_patches, patch_mix, sequential_reference
    t = time.perf_counter()
    scenes = make_patches(patch_mix(count=64))
    log(f"64 patches generated in {time.perf_counter() - t:.1f} s")
    cull = args.cull != "none" and args.mode == "split"
    prog = MultiPatchProgram(scenes, mode=args.mode, merge_interp="kaiser8",
                             persistent=args.persistent, cull=cull)
    t = time.perf_counter()
    tuned_cfg, tune_report, tune_cached = prog.autotune()
    best, pulse_block = tuned_cfg["leaf"]
    log(f"autotuned in {time.perf_counter() - t:.1f} s: {json.dumps(tuned_cfg)}")
    prog.capture_graph()
    for _ in range(args.warmup):
        prog.replay()
    synchronize()
    a, b = Event(), Event()
    a.record()
    for _ in range(args.steps):
        prog.replay()
    b.record()
    synchronize()
    ms_batch = elapsed_ms(a, b) / args.steps
    batch_out = [o.get() for o in prog.outputs]
    lat_p50 = lat_p99 = ms_batch
    if args.persistent:
        # per-patch completion inside the one launch (device wall clock), over
        # a few single replays
        lat = []
        for _ in range(3):
            prog.replay()
            synchronize()
            lat.append(prog.completion_ms())
        lat = np.concatenate(lat)
        lat_p50 = float(np.percentile(lat, 50))
        lat_p99 = float(np.percentile(lat, 99))
    seq_out, isolated, completion, progs = sequential_reference(scenes, mode=args.mode,
                                                                 merge_interp="kaiser8",
                                                                 tile_cols_log2=best, cull=cull,
                                                                 pulse_block=pulse_block)
    same = all(np.array_equal(x, y.get()) for x, y in zip(batch_out, seq_out))
    for _, p_, _, _ in progs:
        p_.close()
    work = prog.work
    seq_total = completion[-1]
    roofline = None
    if not args.persistent:
        # stage times of the batch from events between its launches (eager
        # execution of the same schedule the graph holds)
        timer = StageTimer()
        timer.mark("start")
        reps = 3
        for _ in range(reps):
            prog.execute(mark=timer.mark)
        synchronize()
        stages = timer.durations()
        merge_ms = float(np.sum(stages["merges"])) / reps
        merge_px = int(sum(prog.merge_pixels))
        achieved = merge_px * OPS_MERGE_KAISER8 / (merge_ms * 1e-3) / 1e12
        roofline = {
            "bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": None,
            "kernel": "batched_merge_tiled_kernel<kaiser8>",
            "ops_per_unit": OPS_MERGE_KAISER8, "unit_of_work": "parent pixel",
            "launches_per_step": len(prog.merge_tables),
            "parent_pixels_per_batch": merge_px,
            "stages_ms_per_batch": {k: round(float(np.sum(v)) / reps, 4)
                                    for k, v in stages.items()},
            "note": "dominant stage of the batch (merge launches of every height); events "
                    "between the launches of eager executions"}
    groups = None
    if not args.persistent and args.mode == "split" and args.latency_groups != "none":
        groups = latency_groups(scenes, batch_out, (best, pulse_block), cull, args,
                                tuple(float(f) for f in args.latency_groups.split(",")))
    cpu = None if args.no_cpu else cpu_baseline_multipatch(scenes)
    return {
        "ms_per_step": ms_batch,
        "value": work / (ms_batch * 1e-3) / 1e9,
        "multipatch": {
            "patches": len(scenes), "pixel_pulse_products": work,
            "batched_graph_ms": round(ms_batch, 3),
            "batched_latency_p50_ms": round(lat_p50, 3), "batched_latency_p99_ms": round(lat_p99, 3),
            "batched_latency_source": ("per-patch completion stamps inside the persistent launch"
                                       if args.persistent else
                                       "every patch completes with the batch"),
            "sequential_total_ms": round(seq_total, 3),
            "sequential_latency_p50_ms": round(float(np.percentile(completion, 50)), 3),
            "sequential_latency_p99_ms": round(float(np.percentile(completion, 99)), 3),
            "sequential_isolated_p50_ms": round(float(np.percentile(isolated, 50)), 3),
            "sequential_isolated_p99_ms": round(float(np.percentile(isolated, 99)), 3),
            "sequential_gpix_pulse_per_s": round(work / (seq_total * 1e-3) / 1e9, 2),
            "batched_equals_sequential_bitwise": bool(same),
            "autotune": autotune_summary(tuned_cfg, tune_report, tune_cached,
                                         {"leaf": [6, 32], "merge_shapes": [4, 2176]}),
            "launches_per_batch": prog.launch_count,
            "latency_groups": groups},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "config": {"workload": "c5_multipatch", "patches": len(scenes), "mode": args.mode,
                   "merge_interp": "kaiser8", "graph": True, "persistent": bool(args.persistent),
                   "cull": "grid" if cull else "none"},
    }


def latency_groups(scenes, batch_out, leaf, cull, args, fractions=(0.25, 0.5, 0.75)):
    """The same batch served in latency groups (GroupedMultiPatch: patches by
    work, cut at ``fractions`` of the count, one group after another in one
    HIP graph) with the main batch's leaf shapes, so the outputs are
    the main batch's bit for bit; per-patch completion from events after
    each group of eager executions."""
    from bp_project_amd.device import Event, elapsed_ms, synchronize
    from bp_project_amd.multipatch import GroupedMultiPatch
    gp = GroupedMultiPatch(scenes, fractions=fractions, mode=args.mode, merge_interp="kaiser8",
                           cull=cull, tile_cols_log2=int(leaf[0]), pulse_block=leaf[1])
    try:
        tuned = gp.autotune(leaf=leaf)
        gp.capture_graph()
        for _ in range(max(1, args.warmup)):
            gp.replay()
        synchronize()
        a, b = Event(), Event()
        a.record()
        for _ in range(args.steps):
            gp.replay()
        b.record()
        synchronize()
        ms = elapsed_ms(a, b) / args.steps
        lat = np.concatenate([gp.completion_ms() for _ in range(3)])
        same = all(np.array_equal(o.get(), x) for o, x in zip(gp.outputs, batch_out))
        work = gp.work
        rec = {"fractions": list(fractions), "group_sizes": [len(g) for g in gp.groups],
               "graph_ms": round(ms, 3), "gpix_pulse_per_s": round(work / (ms * 1e-3) / 1e9, 2),
               "latency_p50_ms": round(float(np.percentile(lat, 50)), 3),
               "latency_p99_ms": round(float(np.percentile(lat, 99)), 3),
               "group_completion_ms": [round(float(np.median(lat[np.asarray(g)])), 3)
                                       for g in gp.groups],
               "merge_shapes": [cfg["merge_shapes"] for cfg, _, _ in tuned],
               "launches_per_batch": gp.launch_count,
               "equals_batch_bitwise": bool(same),
               "latency_source": "events after each group's launches, eager executions"}
    finally:
        gp.close()
    log(f"latency groups {rec['group_sizes']}: {rec['graph_ms']} ms per batch, "
        f"p50 {rec['latency_p50_ms']} ms, p99 {rec['latency_p99_ms']} ms")
    return rec


def cpu_baseline_multipatch(scenes, budget_s=10.0):
    """The oracle's FFBP + Cartesian evaluation patch after patch (the
    sequential per-patch loop) over a seeded subset of the batch, until about
    ``budget_s`` of CPU work; effective pixel*pulse/s."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    from bp_project_amd.scenes import polar_extent_for_grid
    order = np.random.default_rng(20260921).permutation(len(scenes))
    work = seconds = 0.0
    done = []
    for i in order:
        s = scenes[int(i)]
        t = time.perf_counter()
        extent = polar_extent_for_grid(s.grid, s.pos.mean(axis=0), margin=2.0)
        res, _ = oracle.form_ffbp(s.data, s.pos, s.r_start, s.data_dr, s.fc, extent, side=1,
                                  leaf_size=16, range_spacing=0.5, angular_oversample=2.5,
                                  weight=s.weight, phase=s.phase, policy="parity",
                                  leaf_interp="lagrange4", merge_interp="kaiser8")
        oracle.evaluate_image(res.image, res.origin, res.r0, res.dr, res.u0, res.du,
                              s.grid.points(), s.k, interp="kaiser8")
        seconds += time.perf_counter() - t
        work += s.grid.nx * s.grid.ny * s.pulses
        done.append(int(i))
        if seconds >= budget_s:
            break
    return {"value": round(work / seconds / 1e9, 4), "unit": "Gpix·pulse/s",
            "cores": oracle.num_threads(), "kind": "port", "seconds": round(seconds, 2),
            "sample": f"FP64 oracle FFBP + evaluate_image of {len(done)} of the {len(scenes)} "
                      f"patches (seeded order {done[:8]}...), one after another; effective "
                      "output pixel*pulse/s"}


def run_range_compress(args):
    """Range compression of an 8192-pulse x 8192-sample raw block (the C2 cube
    before compression): 1 us, 200 MHz down-chirp echoes in noise at 400 MHz,
    no oversampling.  Roofline: HBM, compulsory 16 B per raw sample (complex64
    read and written once)."""
    from bp_project_amd.device import DeviceArray, Event, elapsed_ms, synchronize
    from bp_project_amd.sar_io import fft_length, range_compress_device, transform_length
    pulses, samples, tp = 8192, 8192, 1e-6
    rng = np.random.default_rng(20260919)
    raw = (rng.standard_normal((pulses, samples), np.float32)
           + 1j * rng.standard_normal((pulses, samples), np.float32)).astype(np.complex64)
    d = DeviceArray.from_host(raw)
    out = DeviceArray.empty((pulses, samples))
    nfft = fft_length(samples, tp)
    for _ in range(args.warmup):
        range_compress_device(d, tp, out=out)
    synchronize()
    a, b = Event(), Event()
    a.record()
    for _ in range(args.steps):
        range_compress_device(d, tp, out=out)
    b.record()
    synchronize()
    ms = elapsed_ms(a, b) / args.steps
    bytes_ = 16.0 * pulses * samples
    achieved = bytes_ / (ms * 1e-3) / 1e9
    return {"ms_per_step": ms, "value": pulses * samples / (ms * 1e-3) / 1e6,
            "unit": "Msample/s",
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                         "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                         "traffic": None, "kernel": "range compression pipeline",
                         "unit_of_work": "raw sample", "bytes_per_unit": 16,
                         "note": "compulsory bytes of the whole pipeline (mean/pad, forward "
                                 "FFT, spectrum product, inverse FFT, crop) over its time"},
            "config": {"workload": "range_compress", "pulses": pulses, "samples": samples,
                       "tp_s": tp, "fs_hz": 400e6, "bandwidth_hz": 200e6, "nfft": nfft,
                       "transform_length": transform_length(samples, tp),
                       "oversample": 1,
                       "fft": "fused LDS kernel, one workgroup per pulse (complex64)"}}


def cpu_baseline_range_compress():
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    rng = np.random.default_rng(1)
    rows, samples = 256, 8192
    raw = (rng.standard_normal((rows, samples)) + 1j * rng.standard_normal((rows, samples)))
    raw = raw.astype(np.complex64)
    # repeat the block until ~10 s of CPU work
    t = time.perf_counter()
    reps = 0
    while reps == 0 or time.perf_counter() - t < 10.0:
        oracle.range_compress(raw)
        reps += 1
    dt = time.perf_counter() - t
    rows *= reps
    return {"value": round(rows * samples / dt / 1e6, 3), "unit": "Msample/s", "cores": 1,
            "kind": "port", "seconds": round(dt, 2),
            "sample": f"FP64 oracle (scipy.fft, one worker) range compression of {rows} pulses "
                      f"x {samples} samples"}


def cpu_baseline_direct(scene, name, budget_s=10.0):
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    g = scene.grid
    interp = "linear" if name == "c1_direct_small" else "lagrange4"

    def run(rows):
        pts = g.points()[: rows * g.nx]
        t = time.perf_counter()
        oracle.backproject_points(scene.data, scene.pos, pts, scene.r_start, scene.data_dr,
                                  scene.k, scene.weight, scene.phase, interp=interp)
        return len(pts), time.perf_counter() - t

    # a probe of 2 grid rows sizes a sample of ~budget_s of CPU work (whole
    # grids repeated when the grid is smaller than that)
    n, dt = run(2)
    rows = min(g.ny, max(1, int(2 * budget_s / max(dt, 1e-3))))
    total_px = total_s = 0.0
    reps = 0
    while reps == 0 or total_s < 0.8 * budget_s:
        n, dt = run(rows)
        total_px += n
        total_s += dt
        reps += 1
    return {"value": round(total_px * scene.pulses / total_s / 1e9, 4), "unit": "Gpix·pulse/s",
            "cores": oracle.num_threads(), "kind": "port", "seconds": round(total_s, 2),
            "sample": f"FP64 oracle direct BP of {rows} grid rows x {reps} ({int(total_px)} "
                      f"pixels) x {scene.pulses} pulses"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the multi-GPU sharding of this path is a design note, not built)",
              file=sys.stderr)
        sys.exit(2)
    from bp_project_amd.device import device_count, device_info
    if device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    from bp_project_amd.scenes import workload_scene
    t = time.perf_counter()
    scene = (None if args.workload in ("c5_multipatch", "range_compress")
             else workload_scene(args.workload))
    if scene is not None:
        log(f"scene {args.workload}: {scene.data.shape}, {len(scene.targets)} targets, "
            f"{time.perf_counter() - t:.1f} s")
    log("device:", json.dumps(device_info(0)))
    if args.workload == "c2_ffbp_full":
        res = run_ffbp(args, scene)
        cpu = None if args.no_cpu else cpu_baseline_ffbp(args, scene)
    elif args.workload == "c4_autofocus":
        res = run_autofocus(args, scene)
        cpu = None if args.no_cpu else cpu_baseline_autofocus(scene)
    elif args.workload == "c5_multipatch":
        res = run_multipatch(args)
        cpu = res.get("cpu_baseline")
    elif args.workload == "range_compress":
        res = run_range_compress(args)
        cpu = None if args.no_cpu else cpu_baseline_range_compress()
    else:
        res = run_direct(args, scene, args.workload)
        cpu = None if args.no_cpu else cpu_baseline_direct(scene, args.workload)
    if cpu:
        log("cpu baseline:", json.dumps(cpu))
    line = {
        "metric": METRIC, "value": round(res["value"], 3), "unit": res.get("unit", "Gpix·pulse/s"),
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(res["ms_per_step"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": res["config"], "roofline": attach_measured(res["roofline"]), "cpu_baseline": cpu,
    }
    if "quality" in res:
        line["quality"] = res["quality"]
    for extra in ("timing", "autofocus", "multipatch", "cull", "autotune"):
        if extra in res:
            line[extra] = res[extra]
    if "precision_study" in res:
        line["precision_study"] = res["precision_study"]
    if "stages" in res:
        line["stages_ms_per_step"] = {k: round(v["ms_per_step"], 4) for k, v in res["stages"].items()}
    if args.workload == "c2_ffbp_full" and not args.no_others:
        del scene, res
        line["other_workloads"] = other_workloads(args)
    print(json.dumps(line), flush=True)


def other_workloads(args):
    """The other stripmap workloads, each measured in this same run with
    fewer steps (C1 and C3 direct BP, C4 autofocus, C5 batched patches),
    with their roofline fraction and a shorter CPU baseline (~4 s each):
    compact records of what ``--workload NAME`` reports in full."""
    from bp_project_amd.scenes import workload_scene
    out = {}
    sub = argparse.Namespace(**vars(args))
    sub.steps, sub.warmup = 5, 1
    for name in ("c1_direct_small", "c3_direct_precision", "c4_autofocus", "c5_multipatch"):
        t0 = time.perf_counter()
        scene = None if name == "c5_multipatch" else workload_scene(name)
        sub.workload = name
        sub.precision_study = name == "c3_direct_precision"
        if name == "c4_autofocus":
            res = run_autofocus(sub, scene)
            cpu = None if args.no_cpu else cpu_baseline_autofocus(scene, budget_s=4.0)
        elif name == "c5_multipatch":
            res = run_multipatch(sub)
            cpu = res.get("cpu_baseline")
        else:
            res = run_direct(sub, scene, name)
            cpu = None if args.no_cpu else cpu_baseline_direct(scene, name, budget_s=4.0)
        rec = {"value": round(res["value"], 3), "unit": res.get("unit", "Gpix·pulse/s"),
               "ms_per_step": round(res["ms_per_step"], 4), "steps": sub.steps,
               "config": res["config"]}
        if res.get("roofline"):
            roof = attach_measured(res["roofline"])
            rec["roofline"] = {k: roof[k] for k in ("bound", "achieved", "peak", "unit", "frac",
                                                    "frac_of_measured_fp32", "kernel")
                               if k in roof}
        if cpu:
            rec["cpu_baseline"] = {k: cpu[k] for k in ("value", "unit", "cores", "kind",
                                                       "sample")}
        for key in ("quality", "precision_study", "multipatch"):
            if key in res:
                rec[key] = res[key]
        if "autofocus" in res:
            af = res["autofocus"]
            rec["autofocus"] = {k: af[k] for k in (
                "ms_per_gradient_step", "entropy_initial", "entropy_final",
                "fd_fp64_gradient_step_ms", "speedup_vs_fd_fp64_per_step",
                "equal_entropy_reduced_scene")}
        rec["seconds"] = round(time.perf_counter() - t0, 1)
        log(f"{name}:", json.dumps(rec)[:600])
        out[name] = rec
        del scene, res
    return out


if __name__ == "__main__":
    main()
