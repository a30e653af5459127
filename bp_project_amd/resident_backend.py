"""Preallocated, asynchronous BP programs with HIP graph capture.

Counterpart of fast_bp/resident_backend.py:24-319 with the same constructor,
``execute`` / ``capture_graph`` / ``replay`` / ``close``, ``output``,
``metadata``, ``launch_count``, ``operations`` and ``_schedule``.

All input arrays stay owned by the caller.  Construction allocates every
workspace buffer; ``execute`` and ``replay`` allocate nothing, never
synchronise and record no events: time them with events on the current
stream.  The direct program clears its accumulator with a fill kernel (not a
memset node), so the captured graph replays correctly.  Outputs are
overwritten by the next execution; a program is not safe for concurrent
execution because its workspace is shared between invocations.

A schedule entry is ``(name, launcher, grid, block, args)``; calling
``launcher(grid, block, args)`` enqueues the C-ABI call ``name`` with
``args`` on the current stream.  ``grid``/``block`` describe the launch for
inspection only (the library picks its own launch shape).
"""
from __future__ import annotations

from collections import Counter

import numpy as np

from . import _lib
from .device import DeviceArray, Graph, Stream, current_stream, stream_handle
from .gpu_backend import ALL_MODES, MODES, check_mode, geo_array, leaf_tile_width, origin_reference  # noqa: F401

_vp = _lib.c_void_p


class _Launcher:
    """Enqueue one C-ABI call on the current stream."""

    __slots__ = ("cname",)

    def __init__(self, cname):
        self.cname = cname

    def __call__(self, grid, block, args):
        _lib.call(self.cname, *args, stream_handle())

    def __repr__(self):
        return f"<launch {self.cname}>"


class ResidentProgram:
    """A reusable direct-BP or FFBP schedule with fixed inputs and grid."""

    def __init__(self, backend, plan, device, mode, range_start, data_dr, k,
                 algorithm="ffbp", variant="original", chunk_pulses=64):
        check_mode(mode)
        if algorithm not in ("direct", "ffbp"):
            raise ValueError("algorithm must be direct or ffbp")
        if variant not in ("original", "fused"):
            raise ValueError("variant must be original or fused")
        if algorithm == "direct" and variant != "original":
            raise ValueError("fused variant is only defined for FFBP")
        if chunk_pulses <= 0:
            raise ValueError("chunk_pulses must be positive")
        self.backend = backend
        self.plan = plan
        self.device = device
        self.mode = mode
        self.algorithm = algorithm
        self.variant = variant
        self.range_start = float(range_start)
        self.data_dr = float(data_dr)
        self.k = float(k)
        self.chunk_pulses = int(chunk_pulses)
        self._schedule = []
        self.operations = []
        self._buffers = []
        self._geo = []          # host geometry rows kept alive for the schedule
        self._graph = None
        self._graph_stream = None
        self._closed = False
        self._accumulator = None
        self.output = None
        if algorithm == "direct":
            self._build_direct()
        else:
            root_image = self._build_node(plan.root)
            flat = self._new_buffer(plan.nr * plan.nu)
            self._enqueue("bpmi_crop", plan.nr * plan.nu, (
                _vp(root_image.ptr), plan.root.nu, plan.crop_r[0], plan.crop_u[0], plan.nr,
                plan.nu, _vp(flat.ptr)), name="crop_float2")
            self.output = flat.reshape(plan.shape)
        self.launch_count = len(self._schedule)
        counts = Counter(entry[0] for entry in self._schedule)
        workspace_bytes = sum(int(a.nbytes) for a in self._buffers)
        self.metadata = {
            "algorithm": algorithm,
            "variant": variant,
            "mode": mode,
            "shape": [int(x) for x in plan.shape],
            "block_size": int(backend.block_size),
            "chunk_pulses": self.chunk_pulses if algorithm == "direct" else None,
            "kernel_launch_count": self.launch_count,
            "kernel_counts": dict(counts),
            "memset_count": 0,
            "zero_fill_count": int(algorithm == "direct"),
            "device_operation_count": self.launch_count + int(algorithm == "direct"),
            "workspace_bytes": workspace_bytes,
            "workspace_buffer_count": len(self._buffers),
            "origin_table_bytes": int(np.asarray(plan.origins).nbytes),
            "total_owned_device_bytes": workspace_bytes,
            "workspace_policy": "one device buffer per node image, all made at construction",
            "input_data_shared": True,
            "output_dtype": "complex64",
            "zeroing_in_execute": algorithm == "direct",
            "graph_supported": None,
            "graph_error": None,
            "arithmetic": f"libbpmi {mode} mode",
            "fusion": ("two child regrids rounded to complex64 then FP32 add"
                       if variant == "fused" else None),
        }

    # -- schedule construction ----------------------------------------------------
    def _new_buffer(self, count, dtype=np.complex64):
        array = DeviceArray.empty(int(count), dtype)
        self._buffers.append(array)
        return array

    def _keep(self, host):
        host = np.ascontiguousarray(host)
        self._geo.append(host)
        return host.ctypes.data_as(_vp)

    def _enqueue(self, cname, count, args, name=None, launcher=None):
        count = int(count)
        if count <= 0:
            raise ValueError("empty image grids are unsupported")
        block = self.backend.block_size
        grid = ((count + block - 1) // block,)
        launcher = launcher or _Launcher(cname)
        label = name or cname
        self._schedule.append((label, launcher, grid, (block,), args))
        self.operations.append({"name": label, "count": count, "kernel": launcher,
                                "grid": grid, "block": (block,), "args": args})

    def _interp(self):
        return _lib.interp_code(self.backend.interp)

    def _merge_interp(self):
        return _lib.interp_code(self.backend.merge_interp)

    def _build_direct(self):
        p, d = self.plan, self.device
        count = p.nr * p.nu
        self._accumulator = self._new_buffer(
            count, np.complex128 if self.mode == "fp64" else np.complex64)
        flat = self._new_buffer(count) if self.mode == "fp64" else self._accumulator
        pulses = int(d.data.shape[0])
        geo = self._keep(geo_array(p))
        flags = _lib.ACCUMULATE | (_lib.OUT_C128 if self.mode == "fp64" else 0)
        oref = 0
        if self.mode == "split":
            oref = origin_reference(d._host_position, 0, pulses, p.r0)
        for begin in range(0, pulses, self.chunk_pulses):
            end = min(pulses, begin + self.chunk_pulses)
            self._enqueue("bpmi_bp_polar_ex", count, (
                _vp(d.data.ptr), d.data.shape[1], _vp(d.pos64.ptr), _vp(d.weight64.ptr),
                _vp(d.phase64.ptr), begin, end, geo, p.nr, p.nu, p.side, self.range_start,
                self.data_dr, self.k, _lib.mode_code(self.mode), self._interp(), flags,
                _vp(self._accumulator.ptr)), name=f"direct_chunk_{self.mode}",
                launcher=_ExLauncher("bpmi_bp_polar_ex", oref, -1))
        if self.mode == "fp64":
            self._enqueue("bpmi_cast_c128_c64", count, (
                _vp(self._accumulator.ptr), _vp(flat.ptr), count), name="cast_double2_to_float2")
        self.output = flat.reshape(p.shape)

    def _leaf(self, node):
        d = self.device
        output = self._new_buffer(node.nr * node.nu)
        geo = self._keep(geo_array(node))
        args = (_vp(d.data.ptr), d.data.shape[1], _vp(d.pos64.ptr), _vp(d.weight64.ptr),
                _vp(d.phase64.ptr), node.a, node.b, geo, node.nr, node.nu, node.side,
                self.range_start, self.data_dr, self.k, _lib.mode_code(self.mode),
                self._interp(), 0, _vp(output.ptr))
        oref, tc = 0, -1
        if self.mode == "split":
            oref = origin_reference(d._host_position, node.a, node.b, node.r0)
            if getattr(self, "_leaf_tc", None) is None:
                self._leaf_tc = leaf_tile_width(self.plan, d._host_position, self.data_dr)
            tc = self._leaf_tc
        self._enqueue("bpmi_bp_polar_ex", node.nr * node.nu, args, name=f"bp_polar_{self.mode}",
                      launcher=_ExLauncher("bpmi_bp_polar_ex", oref, tc))
        return output

    def _regrid(self, image, source, destination, accumulate_into=None):
        output = accumulate_into if accumulate_into is not None else \
            self._new_buffer(destination.nr * destination.nu)
        self._enqueue("bpmi_regrid", destination.nr * destination.nu, (
            _vp(image.ptr), self._keep(geo_array(source)), source.nr, source.nu,
            self._keep(geo_array(destination)), destination.nr, destination.nu, destination.side,
            self.k, _lib.mode_code(self.mode), self._merge_interp(),
            _lib.ACCUMULATE if accumulate_into is not None else 0, _vp(output.ptr)),
            name=f"regrid_{self.mode}")
        return output

    def _build_node(self, node):
        if node.is_leaf:
            return self._leaf(node)
        left_child = self._build_node(node.left)
        if self.variant == "original":
            # the reference's depth-first order: regrid left, recurse right,
            # regrid right, add (fast_bp/resident_backend.py:203-211)
            left = self._regrid(left_child, node.left, node)
            right_child = self._build_node(node.right)
            right = self._regrid(right_child, node.right, node)
            self._enqueue("bpmi_add_inplace", node.nr * node.nu, (
                _vp(left.ptr), _vp(right.ptr), node.nr * node.nu), name="add_float2_inplace")
            return left
        right_child = self._build_node(node.right)
        output = self._new_buffer(node.nr * node.nu)
        self._enqueue("bpmi_merge2", node.nr * node.nu, (
            _vp(left_child.ptr), self._keep(geo_array(node.left)), node.left.nr, node.left.nu,
            _vp(right_child.ptr), self._keep(geo_array(node.right)), node.right.nr,
            node.right.nu, self._keep(geo_array(node)), node.nr, node.nu, node.side, self.k,
            _lib.mode_code(self.mode), self._merge_interp(), _vp(output.ptr)),
            name=f"merge_regrid_{self.mode}")
        return output

    # -- execution -------------------------------------------------------------------
    def execute(self):
        """Enqueue one complete image computation on the current stream."""
        if self._closed:
            raise RuntimeError("program is closed")
        if self._accumulator is not None:
            self._accumulator.fill_zero()
        for _, launcher, grid, block, args in self._schedule:
            launcher(grid, block, args)
        return self.output

    def capture_graph(self, warmup=True):
        """Record one ``execute`` as a HIP graph on a private stream (after an
        untimed warm-up execution there).  True on success; on failure False,
        with the exception text in ``metadata["graph_error"]``.  Set-up code:
        it synchronises the device, which ``execute`` and ``replay`` never do.
        """
        if self._closed:
            raise RuntimeError("program is closed")
        if self._graph is not None:
            return True
        try:
            stream = Stream()
            current_stream().synchronize()
            if warmup:
                with stream:
                    self.execute()
                stream.synchronize()
            graph = Graph.capture(stream, self.execute)
            self._graph = graph
            self._graph_stream = stream
            self.metadata["graph_supported"] = True
            self.metadata["graph_error"] = None
            return True
        except Exception as exc:  # recorded, as the reference does
            self._graph = None
            self.metadata["graph_supported"] = False
            self.metadata["graph_error"] = f"{type(exc).__name__}: {exc}"
            return False

    def replay(self):
        """Launch the recorded graph on the current stream and return the
        output array (asynchronous; the caller orders and waits)."""
        if self._closed:
            raise RuntimeError("program is closed")
        if self._graph is None:
            raise RuntimeError("capture_graph() must succeed before replay()")
        self._graph.launch()
        return self.output

    def close(self):
        """Release the graph, stream and buffers; the caller first waits for
        any work that still uses them."""
        if self._graph is not None:
            self._graph.close()
        self._graph = None
        if self._graph_stream is not None:
            self._graph_stream.close()
        self._graph_stream = None
        self._schedule.clear()
        self.operations.clear()
        self._buffers.clear()
        self._geo.clear()
        self._accumulator = None
        self.output = None
        self.device = None
        self._closed = True


class _ExLauncher:
    """Launcher for bpmi_bp_polar_ex: appends the stream, origin_ref and tile width."""

    __slots__ = ("cname", "origin_ref", "tile_cols_log2")

    def __init__(self, cname, origin_ref, tile_cols_log2):
        self.cname = cname
        self.origin_ref = int(origin_ref)
        self.tile_cols_log2 = int(tile_cols_log2)

    def __call__(self, grid, block, args):
        _lib.call(self.cname, *args, stream_handle(), self.origin_ref, self.tile_cols_log2)

    def __repr__(self):
        return f"<launch {self.cname} ref={self.origin_ref}>"
