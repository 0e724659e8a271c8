"""Device buffers over the C ABI (no torch required) and torch interop."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi
from ._capi import P, check, lib


class DeviceBuffer:
    """A hipMalloc'd region owned by a context; freed on close()/GC."""

    def __init__(self, ctx: _capi.Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self._ptr = P()
        check(lib().csk_malloc(ctx.ptr, ctypes.byref(self._ptr), self.nbytes), "csk_malloc")

    @property
    def ptr(self) -> int:
        return self._ptr.value or 0

    @classmethod
    def from_array(cls, ctx, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        buf = cls(ctx, arr.nbytes)
        buf.upload(arr)
        return buf

    def upload(self, arr: np.ndarray, stream=None) -> None:
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("upload larger than buffer")
        check(lib().csk_memcpy_h2d(self.ctx.ptr, self._ptr, arr.ctypes.data_as(P), arr.nbytes, stream),
              "csk_memcpy_h2d")
        check(lib().csk_stream_sync(self.ctx.ptr, stream), "csk_stream_sync")

    def download(self, shape, dtype=np.uint32, stream=None) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("download larger than buffer")
        check(lib().csk_memcpy_d2h(self.ctx.ptr, out.ctypes.data_as(P), self._ptr, out.nbytes, stream),
              "csk_memcpy_d2h")
        return out

    def close(self) -> None:
        if self._ptr:
            lib().csk_free(self.ctx.ptr, self._ptr)
            self._ptr = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def pinned_empty(shape, dtype=np.uint32, ctx=None) -> np.ndarray:
    """An uninitialised numpy array in page-locked host memory
    (csk_host_alloc), freed when the array and its views are gone.  Host
    buffers that are copied to the device often (a prover's witness) move at
    the link's full rate from it."""
    import weakref
    ctx = ctx or _capi.default_context()
    dt = np.dtype(dtype)
    count = int(np.prod(shape, dtype=np.int64))
    nbytes = count * dt.itemsize
    p = P()
    check(lib().csk_host_alloc(ctx.ptr, ctypes.byref(p), nbytes), "csk_host_alloc")
    raw = (ctypes.c_uint8 * max(nbytes, 1)).from_address(p.value)
    weakref.finalize(raw, _host_free, ctx, p.value)  # holds the context until then
    return np.frombuffer(raw, dtype=np.uint8, count=nbytes).view(dt).reshape(shape)


def _host_free(ctx, addr: int) -> None:
    lib().csk_host_free(ctx.ptr, addr)


def dptr(x) -> int:
    """Device address of a DeviceBuffer, a torch tensor or a raw int."""
    if isinstance(x, DeviceBuffer):
        return x.ptr
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        if not x.is_cuda:
            raise ValueError("tensor is not on the GPU")
        if not x.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return x.data_ptr()
    raise TypeError("not a device buffer: %r" % type(x))


class Stream:
    """A non-blocking HIP stream at the default priority."""

    def __init__(self, ctx: _capi.Context):
        self.ctx = ctx
        self._ptr = P()
        check(lib().csk_stream_create(ctx.ptr, ctypes.byref(self._ptr)), "csk_stream_create")

    @property
    def ptr(self):
        return self._ptr

    def sync(self) -> None:
        check(lib().csk_stream_sync(self.ctx.ptr, self._ptr), "csk_stream_sync")

    def wait(self, event: "Event") -> None:
        check(lib().csk_stream_wait_event(self.ctx.ptr, self._ptr, event.ptr), "csk_stream_wait_event")

    def close(self) -> None:
        if self._ptr:
            lib().csk_stream_destroy(self.ctx.ptr, self._ptr)
            self._ptr = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    """A HIP event (timing enabled)."""

    def __init__(self, ctx: _capi.Context):
        self.ctx = ctx
        self._ptr = P()
        check(lib().csk_event_create(ctx.ptr, ctypes.byref(self._ptr)), "csk_event_create")

    @property
    def ptr(self):
        return self._ptr

    def record(self, stream=None) -> None:
        st = stream.ptr if isinstance(stream, Stream) else stream
        check(lib().csk_event_record(self.ctx.ptr, self._ptr, st), "csk_event_record")

    def done(self) -> bool:
        rc = lib().csk_event_query(self.ctx.ptr, self._ptr)
        if rc < 0:
            check(rc, "csk_event_query")
        return rc == 1

    def sync(self) -> None:
        check(lib().csk_event_sync(self.ctx.ptr, self._ptr), "csk_event_sync")

    def elapsed_ms(self, end: "Event") -> float:
        ms = ctypes.c_float(0)
        check(lib().csk_event_elapsed(self.ctx.ptr, self._ptr, end.ptr, ctypes.byref(ms)), "csk_event_elapsed")
        return float(ms.value)

    def close(self) -> None:
        if self._ptr:
            lib().csk_event_destroy(self.ctx.ptr, self._ptr)
            self._ptr = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
