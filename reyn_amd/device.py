"""Device context and buffers over the C-ABI (no torch, no second HIP runtime).

``Context`` owns one ``reyn_ctx`` (a HIP stream on one device). ``DeviceArray``
is a typed device allocation made through ``reyn_malloc``; it mirrors the
storage role of the reference's ``DeviceBuffer<T>`` (src/buffer.cuh:15-269).
"""

from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import capi


class Context:
    def __init__(self, device: int = 0):
        self._lib = capi.lib()
        h = C.c_void_p()
        rc = self._lib.reyn_ctx_create(int(device), C.byref(h))
        if rc != capi.OK:
            raise capi.ReynError(rc, f"cannot create a context on HIP device {device}")
        self.h = h
        self.device = device
        # objects holding library resources of this context; they are released
        # before the context itself, whatever order Python finalises them in
        self._children = weakref.WeakSet()

    def adopt(self, obj) -> None:
        self._children.add(obj)

    # -- errors ------------------------------------------------------------
    def last_error(self) -> str:
        msg = self._lib.reyn_last_error(self.h)
        return msg.decode() if msg else ""

    def check(self, rc: int, what: str = "") -> int:
        if rc != capi.OK:
            raise capi.ReynError(rc, f"{what}: {self.last_error()}" if what else self.last_error())
        return rc

    # -- lifetime ------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "h", None):
            for child in list(self._children):
                child.release()
            self._lib.reyn_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self) -> None:
        self.check(self._lib.reyn_ctx_synchronize(self.h), "synchronize")

    # -- timing ----------------------------------------------------------------
    def timer_start(self) -> None:
        self.check(self._lib.reyn_timer_start(self.h), "timer_start")

    def timer_stop(self) -> float:
        ms = C.c_float()
        self.check(self._lib.reyn_timer_stop(self.h, C.byref(ms)), "timer_stop")
        return float(ms.value)

    # -- allocation ----------------------------------------------------------------
    def empty(self, n: int, dtype=np.float32) -> "DeviceArray":
        return DeviceArray(self, n, dtype)

    def array(self, host) -> "DeviceArray":
        host = np.ascontiguousarray(host)
        a = DeviceArray(self, host.size, host.dtype)
        a.upload(host)
        return a

    def full(self, n: int, value, dtype=np.float32) -> "DeviceArray":
        a = DeviceArray(self, n, dtype)
        a.fill(value)
        return a


class DeviceArray:
    """A 1-D device allocation of ``n`` elements of ``dtype``."""

    def __init__(self, ctx: Context, n: int, dtype=np.float32):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = C.c_void_p()
        ctx.check(ctx._lib.reyn_malloc(ctx.h, C.byref(p), max(self.nbytes, 4)), "malloc")
        self._ptr = p
        ctx.adopt(self)

    @property
    def nbytes(self) -> int:
        return self.n * self.dtype.itemsize

    @property
    def ptr(self) -> int:
        return self._ptr.value

    def upload(self, host) -> "DeviceArray":
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if host.size != self.n:
            raise ValueError(f"upload of {host.size} elements into an array of {self.n}")
        if self.n:
            self.ctx.check(
                self.ctx._lib.reyn_memcpy(self.ctx.h, self._ptr, host.ctypes.data_as(C.c_void_p), self.nbytes, 0),
                "upload",
            )
        return self

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        if self.n:
            self.ctx.check(
                self.ctx._lib.reyn_memcpy(self.ctx.h, out.ctypes.data_as(C.c_void_p), self._ptr, self.nbytes, 1),
                "download",
            )
        return out

    def copy_from(self, other: "DeviceArray") -> None:
        if other.nbytes != self.nbytes:
            raise ValueError("device copy between arrays of different size")
        if self.n:
            self.ctx.check(self.ctx._lib.reyn_memcpy(self.ctx.h, self._ptr, other._ptr, self.nbytes, 2), "copy")

    def fill(self, value) -> None:
        if self.n == 0:
            return
        if self.dtype == np.float32:
            self.ctx.check(self.ctx._lib.reyn_fill_f32(self.ctx.h, self._ptr, float(value), self.n), "fill")
        elif self.dtype == np.uint32:
            self.ctx.check(self.ctx._lib.reyn_fill_u32(self.ctx.h, self._ptr, int(value), self.n), "fill")
        else:
            self.upload(np.full(self.n, value, dtype=self.dtype))

    def free(self) -> None:
        if getattr(self, "_ptr", None) is not None and self._ptr.value and self.ctx.h:
            self.ctx._lib.reyn_free(self.ctx.h, self._ptr)
        self._ptr = C.c_void_p()

    release = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __len__(self) -> int:
        return self.n


def device_count() -> int:
    n = C.c_int()
    capi.lib().reyn_device_count(C.byref(n))
    return int(n.value)
