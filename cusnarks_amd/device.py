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
