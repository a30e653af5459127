"""Device memory, streams, events and graphs over the libbpmi C ABI.

``DeviceArray`` stands in for the ``cupy.ndarray`` the reference backends
return (``cp.ndarray`` cannot exist without CUDA): it has ``shape``,
``dtype``, ``nbytes``, ``reshape``, first-axis slicing (row views),
``data.ptr`` and ``get()``; ``asnumpy()`` copies to the host.  Memory comes
from ``hipMalloc`` through a caching pool (blocks are recycled in stream
order, like CuPy's default pool), so freeing and reallocating inside a
schedule never synchronises the device.

Work is enqueued on the *current stream* (the null stream unless a
``Stream`` context is active), mirroring CuPy's current-stream model.
"""
from __future__ import annotations

import ctypes
import threading
from ctypes import c_float, c_void_p

import numpy as np

from . import _lib

_tls = threading.local()


# ---------------------------------------------------------------------------
# streams, events, graphs
# ---------------------------------------------------------------------------

class Stream:
    """A HIP stream (non-blocking w.r.t. the null stream); context manager
    that makes it the current stream."""

    def __init__(self, handle=None, owned=True):
        if handle is None:
            h = c_void_p()
            _lib.call("bpmi_stream_create", ctypes.byref(h))
            handle = h.value
        self.ptr = handle or 0
        self._owned = owned and bool(handle)

    def synchronize(self):
        _lib.call("bpmi_stream_synchronize", c_void_p(self.ptr or None))

    def __enter__(self):
        stack = getattr(_tls, "streams", None)
        if stack is None:
            stack = _tls.streams = []
        stack.append(self)
        return self

    def __exit__(self, *exc):
        _tls.streams.pop()

    def close(self):
        if self._owned and self.ptr:
            _lib.call("bpmi_stream_destroy", c_void_p(self.ptr))
        self.ptr = 0
        self._owned = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


NULL_STREAM = Stream(handle=0, owned=False)


def current_stream():
    stack = getattr(_tls, "streams", None)
    return stack[-1] if stack else NULL_STREAM


def stream_handle(stream=None):
    s = current_stream() if stream is None else stream
    return c_void_p(s.ptr or None)


class Event:
    def __init__(self):
        h = c_void_p()
        _lib.call("bpmi_event_create", ctypes.byref(h))
        self.ptr = h.value

    def record(self, stream=None):
        _lib.call("bpmi_event_record", c_void_p(self.ptr), stream_handle(stream))

    def synchronize(self):
        _lib.call("bpmi_event_synchronize", c_void_p(self.ptr))

    def __del__(self):
        try:
            if self.ptr:
                _lib.load().bpmi_event_destroy(c_void_p(self.ptr))
        except Exception:
            pass
        self.ptr = 0


def elapsed_ms(start, end):
    ms = c_float()
    _lib.call("bpmi_event_elapsed", ctypes.byref(ms), c_void_p(start.ptr), c_void_p(end.ptr))
    return float(ms.value)


class Graph:
    """An instantiated HIP graph captured from a stream."""

    def __init__(self, handle):
        self.ptr = handle

    @classmethod
    def capture(cls, stream, enqueue):
        _lib.call("bpmi_graph_begin", c_void_p(stream.ptr))
        h = c_void_p()
        try:
            with stream:
                enqueue()
        finally:
            status = _lib.load().bpmi_graph_end(c_void_p(stream.ptr), ctypes.byref(h))
        _lib.check(status, "bpmi_graph_end")
        return cls(h.value)

    def launch(self, stream=None):
        _lib.call("bpmi_graph_launch", c_void_p(self.ptr), stream_handle(stream))

    def close(self):
        if self.ptr:
            _lib.load().bpmi_graph_destroy(c_void_p(self.ptr))
        self.ptr = 0

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def synchronize():
    _lib.call("bpmi_device_synchronize")


# ---------------------------------------------------------------------------
# memory
# ---------------------------------------------------------------------------

class _Pool:
    """Size-bucketed cache of hipMalloc blocks (stream-ordered reuse)."""

    def __init__(self):
        self.free = {}
        self.used = 0
        self.cached = 0
        self.peak = 0
        self.lock = threading.Lock()

    @staticmethod
    def bucket(n):
        if n <= 512:
            return 512
        # round up to 1/8 of the next power of two
        p = 1 << (int(n - 1).bit_length())
        step = max(512, p // 8)
        return -(-n // step) * step

    def alloc(self, nbytes):
        size = self.bucket(max(1, int(nbytes)))
        with self.lock:
            blocks = self.free.get(size)
            if blocks:
                ptr = blocks.pop()
                self.cached -= size
                self.used += size
                self.peak = max(self.peak, self.used)
                return ptr, size
        h = c_void_p()
        status = _lib.load().bpmi_malloc(ctypes.byref(h), size)
        if status != _lib.OK:
            # release cached blocks and retry once
            self.release_cached()
            _lib.call("bpmi_malloc", ctypes.byref(h), size)
        with self.lock:
            self.used += size
            self.peak = max(self.peak, self.used)
        return h.value, size

    def give_back(self, ptr, size):
        with self.lock:
            self.free.setdefault(size, []).append(ptr)
            self.used -= size
            self.cached += size

    def release_cached(self):
        synchronize()
        with self.lock:
            blocks = [p for lst in self.free.values() for p in lst]
            self.free.clear()
            self.cached = 0
        for p in blocks:
            _lib.call("bpmi_free", c_void_p(p))


POOL = _Pool()


class _Allocation:
    __slots__ = ("ptr", "size", "__weakref__")

    def __init__(self, nbytes):
        self.ptr, self.size = POOL.alloc(nbytes)

    def __del__(self):
        try:
            if self.ptr:
                POOL.give_back(self.ptr, self.size)
        except Exception:
            pass
        self.ptr = 0


class _MemPtr:
    __slots__ = ("ptr",)

    def __init__(self, ptr):
        self.ptr = ptr


class DeviceArray:
    """C-contiguous device array (a view onto an owned allocation)."""

    def __init__(self, shape, dtype, alloc=None, offset=0):
        self.shape = tuple(int(x) for x in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        n = self.nbytes
        self._alloc = alloc if alloc is not None else (_Allocation(n) if n else None)
        self._offset = int(offset)

    # -- construction ------------------------------------------------------
    @classmethod
    def empty(cls, shape, dtype=np.complex64):
        return cls(shape, dtype)

    @classmethod
    def zeros(cls, shape, dtype=np.complex64, stream=None):
        a = cls(shape, dtype)
        if a.nbytes:
            _lib.call("bpmi_zero", c_void_p(a.ptr), a.nbytes, stream_handle(stream))
        return a

    @classmethod
    def from_host(cls, host, dtype=None, stream=None):
        host = np.ascontiguousarray(host, dtype=dtype)
        a = cls(host.shape, host.dtype)
        a.set(host, stream=stream)
        return a

    # -- properties ----------------------------------------------------------
    @property
    def size(self):
        return int(np.prod(self.shape, dtype=np.int64)) if self.shape else 1

    @property
    def nbytes(self):
        return self.size * self.dtype.itemsize

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def ptr(self):
        if self._alloc is None:
            return 0
        return self._alloc.ptr + self._offset

    @property
    def data(self):
        return _MemPtr(self.ptr)

    @property
    def __cuda_array_interface__(self):
        typestr = self.dtype.str
        return {"shape": self.shape, "typestr": typestr, "data": (self.ptr, False),
                "version": 3, "strides": None, "stream": None}

    # -- views ---------------------------------------------------------------
    def reshape(self, *shape):
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
            shape = tuple(shape[0])
        shape = list(shape)
        if -1 in shape:
            i = shape.index(-1)
            rest = int(np.prod([s for j, s in enumerate(shape) if j != i]))
            shape[i] = self.size // max(rest, 1)
        if int(np.prod(shape)) != self.size:
            raise ValueError(f"cannot reshape {self.shape} to {tuple(shape)}")
        return DeviceArray(tuple(shape), self.dtype, self._alloc, self._offset)

    def ravel(self):
        return self.reshape(self.size)

    def __getitem__(self, key):
        if isinstance(key, int):
            key = slice(key, key + 1)
            squeeze = True
        else:
            squeeze = False
        if not isinstance(key, slice):
            raise TypeError("DeviceArray supports first-axis integer/slice indexing only")
        start, stop, step = key.indices(self.shape[0])
        if step != 1:
            raise ValueError("DeviceArray slices must be contiguous")
        rows = max(0, stop - start)
        row_bytes = self.nbytes // max(self.shape[0], 1)
        shape = (rows,) + self.shape[1:]
        view = DeviceArray(shape, self.dtype, self._alloc, self._offset + start * row_bytes)
        return view.reshape(self.shape[1:]) if squeeze and self.ndim > 1 else view

    def __len__(self):
        return self.shape[0]

    # -- transfers -------------------------------------------------------------
    def set(self, host, stream=None):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if host.nbytes != self.nbytes:
            raise ValueError("host array size does not match")
        _lib.call("bpmi_memcpy_h2d", c_void_p(self.ptr), host.ctypes.data_as(c_void_p),
                  self.nbytes, stream_handle(stream))
        # the copy reads pageable host memory; wait so the source may be freed
        (current_stream() if stream is None else stream).synchronize()

    def get(self, stream=None):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _lib.call("bpmi_memcpy_d2h", out.ctypes.data_as(c_void_p), c_void_p(self.ptr),
                      self.nbytes, stream_handle(stream))
            (current_stream() if stream is None else stream).synchronize()
        return out

    def copy_from(self, other, stream=None):
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _lib.call("bpmi_memcpy_d2d", c_void_p(self.ptr), c_void_p(other.ptr), self.nbytes,
                  stream_handle(stream))

    def fill_zero(self, stream=None):
        if self.nbytes:
            _lib.call("bpmi_zero", c_void_p(self.ptr), self.nbytes, stream_handle(stream))

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype}, ptr=0x{self.ptr:x})"


def asnumpy(array):
    return array.get() if isinstance(array, DeviceArray) else np.asarray(array)


def to_device(host, dtype=None, stream=None):
    return DeviceArray.from_host(host, dtype=dtype, stream=stream)


def device_info(device=0):
    name = ctypes.create_string_buffer(256)
    arch = ctypes.create_string_buffer(256)
    cu = ctypes.c_int()
    clk = ctypes.c_int()
    mem = ctypes.c_size_t()
    _lib.call("bpmi_device_info", device, name, 256, arch, 256, ctypes.byref(cu),
              ctypes.byref(clk), ctypes.byref(mem))
    ver = ctypes.c_int()
    _lib.call("bpmi_runtime_version", ctypes.byref(ver))
    return {"name": name.value.decode(errors="replace"), "arch": arch.value.decode(),
            "compute_units": cu.value, "clock_khz": clk.value, "memory_bytes": mem.value,
            "hip_runtime_version": ver.value}


def device_count():
    n = ctypes.c_int()
    status = _lib.load().bpmi_device_count(ctypes.byref(n))
    return n.value if status == _lib.OK else 0
