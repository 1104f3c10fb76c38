"""Device memory, streams, timers and graphs over the libumi_hip C ABI.

``DeviceBuffer`` plays the role of the reference's ``MemoryBuffer``/
``SyncedMemory`` pair (src/SyncedMemory.h:5-116, johnjohnlin/UMI) for Python
callers: it owns one device allocation and copies numpy arrays in and out.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import call, lib

H2D, D2H, D2D = 1, 2, 3


def _handle(stream) -> int | None:
    return None if stream is None else stream.handle


class DeviceBuffer:
    """One device allocation of ``nbytes`` bytes (freed on ``free()``/GC)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        call("umi_malloc", C.byref(p), max(self.nbytes, 1))
        self.ptr = p.value or 0

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    @classmethod
    def zeros(cls, nbytes: int) -> "DeviceBuffer":
        b = cls(nbytes)
        call("umi_memset", b.ptr, 0, b.nbytes)
        return b

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError(f"array of {a.nbytes} B does not fit a {self.nbytes} B buffer")
        call("umi_memcpy", self.ptr, a.ctypes.data, a.nbytes, H2D)

    def download(self, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError(f"requested {out.nbytes} B from a {self.nbytes} B buffer")
        call("umi_memcpy", out.ctypes.data, self.ptr, out.nbytes, D2H)
        return out

    def fill(self, byte: int, stream=None) -> None:
        """Every byte = `byte`.  With a stream, a fill kernel ordered on that
        stream (nbytes a multiple of 4); without, a synchronous memset."""
        if stream is None:
            call("umi_memset", self.ptr, int(byte), self.nbytes)
        else:
            if self.nbytes % 4:
                raise ValueError("stream-ordered fill needs a whole number of 32-bit words")
            call("umi_fill_u32", self.ptr, (int(byte) & 0xFF) * 0x01010101, self.nbytes // 4, _handle(stream))

    def free(self) -> None:
        if getattr(self, "ptr", 0):
            lib().umi_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HostBuffer:
    """Pinned host memory (umi_host_alloc) of ``nbytes`` bytes, viewed as numpy."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        call("umi_host_alloc", C.byref(p), max(self.nbytes, 1))
        self.ptr = p.value or 0

    def array(self, shape, dtype) -> np.ndarray:
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        if n > self.nbytes:
            raise ValueError(f"view of {n} B on a {self.nbytes} B buffer")
        buf = (C.c_char * n).from_address(self.ptr)
        return np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)

    def free(self) -> None:
        if getattr(self, "ptr", 0):
            lib().umi_host_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream (default priority)."""

    def __init__(self):
        p = C.c_void_p()
        call("umi_stream_create", C.byref(p))
        self.handle = p.value

    def wait(self, other: "Stream") -> None:
        """Work enqueued here from now on waits for the work enqueued on `other` so far."""
        call("umi_stream_wait", self.handle, other.handle)

    def synchronize(self) -> None:
        call("umi_stream_synchronize", self.handle)

    def destroy(self) -> None:
        if self.handle:
            lib().umi_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Timer:
    """hipEvent pair recorded on a stream; ``ms()`` waits for the stop event."""

    def __init__(self):
        p = C.c_void_p()
        call("umi_timer_create", C.byref(p))
        self.handle = p.value

    def start(self, stream: "Stream | None" = None) -> None:
        call("umi_timer_start", self.handle, _handle(stream))

    def stop(self, stream: "Stream | None" = None) -> None:
        call("umi_timer_stop", self.handle, _handle(stream))

    def ms(self) -> float:
        v = C.c_float()
        call("umi_timer_elapsed_ms", self.handle, C.byref(v))
        return float(v.value)

    def __del__(self):
        try:
            lib().umi_timer_destroy(self.handle)
        except Exception:
            pass


class Graph:
    """Captures the launches issued on ``stream`` between ``begin``/``end``."""

    def __init__(self, stream: Stream):
        self.stream = stream
        self.handle = None

    def begin(self) -> None:
        call("umi_graph_begin", self.stream.handle)

    def end(self) -> None:
        p = C.c_void_p()
        call("umi_graph_end", self.stream.handle, C.byref(p))
        self.handle = p.value

    def launch(self) -> None:
        call("umi_graph_launch", self.handle, self.stream.handle)

    def __del__(self):
        try:
            if self.handle:
                lib().umi_graph_destroy(self.handle)
        except Exception:
            pass


def synchronize() -> None:
    call("umi_synchronize")


def device_count() -> int:
    n = C.c_int(0)
    st = lib().umi_device_count(C.byref(n))
    return n.value if st == 0 else 0


def device_name() -> str:
    buf = C.create_string_buffer(256)
    call("umi_device_name", buf, 256)
    return buf.value.decode()
