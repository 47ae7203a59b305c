"""Host helpers over the C ABI: a device context, device arrays and status
checking.  Memory comes from the library's pool (mbx_alloc); nothing here
touches torch."""
import ctypes as C

import numpy as np

from . import abi, load


class MbxError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"mbx status {status}: {msg}")
        self.status = status


def check(status, lib=None):
    if status != abi.OK:
        lib = lib or load()
        detail = lib.mbx_last_error().decode(errors="replace")
        raise MbxError(status, f"{lib.mbx_strerror(status).decode()}: {detail}")
    return status


def device_count():
    lib = load()
    n = C.c_int(0)
    check(lib.mbx_device_count(C.byref(n)))
    return n.value


class DevArray:
    """A pool allocation on the device with a dtype and length."""

    def __init__(self, dev, ptr, nbytes, dtype=np.uint8, n=None):
        self.dev, self.ptr, self.nbytes = dev, ptr, nbytes
        self.dtype = np.dtype(dtype)
        self.n = n if n is not None else nbytes // self.dtype.itemsize

    def host(self, n=None, dtype=None):
        dtype = np.dtype(dtype or self.dtype)
        n = self.nbytes // dtype.itemsize if n is None else n
        out = np.empty(n, dtype=dtype)
        if n:
            check(self.dev.lib.mbx_memcpy_d2h(self.dev.ctx, out.ctypes.data,
                                              self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            check(self.dev.lib.mbx_free(self.dev.ctx, self.ptr))
            self.ptr = None

    def __int__(self):
        return self.ptr


class Device:
    """mbx_ctx on one GPU."""

    def __init__(self, device=0, pool_bytes=0):
        self.lib = load()
        ctx = C.c_void_p()
        check(self.lib.mbx_init(device, pool_bytes, C.byref(ctx)))
        self.ctx = ctx

    def close(self):
        if self.ctx:
            self.lib.mbx_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def alloc(self, nbytes, dtype=np.uint8, n=None):
        p = C.c_void_p()
        check(self.lib.mbx_alloc(self.ctx, max(int(nbytes), 1), C.byref(p)))
        return DevArray(self, p.value, int(nbytes), dtype, n)

    def zeros(self, n, dtype):
        dtype = np.dtype(dtype)
        a = self.alloc(n * dtype.itemsize, dtype, n)
        nb = n * dtype.itemsize
        check(self.lib.mbx_fill_u32(self.ctx, a.ptr, 0, (nb + 3) // 4)
              if nb % 4 == 0 else self.lib.mbx_memcpy_h2d(
                  self.ctx, a.ptr, np.zeros(nb, np.uint8).ctypes.data, nb))
        return a

    def upload(self, arr, pad_to=None):
        arr = np.ascontiguousarray(arr)
        nb = arr.nbytes if pad_to is None else max(arr.nbytes, pad_to)
        a = self.alloc(nb, arr.dtype, arr.size)
        if nb > arr.nbytes:
            check(self.lib.mbx_fill_u32(self.ctx, a.ptr, 0, (nb + 3) // 4))
        if arr.nbytes:
            check(self.lib.mbx_memcpy_h2d(self.ctx, a.ptr, arr.ctypes.data,
                                          arr.nbytes))
        return a

    def sync(self):
        check(self.lib.mbx_sync(self.ctx))

    def pool_stats(self):
        v = [C.c_size_t() for _ in range(4)]
        check(self.lib.mbx_pool_stats(self.ctx, *[C.byref(x) for x in v]))
        return dict(zip(("used", "capacity", "peak", "nlive"),
                        (x.value for x in v)))


class Timer:
    """HIP events on the context's stream (the stream the kernels run on)."""

    def __init__(self, dev):
        self.dev = dev
        t = C.c_void_p()
        check(dev.lib.mbx_timer_create(dev.ctx, C.byref(t)))
        self.t = t

    def start(self):
        check(self.dev.lib.mbx_timer_start(self.dev.ctx, self.t))

    def stop(self):
        ms = C.c_float()
        check(self.dev.lib.mbx_timer_stop(self.dev.ctx, self.t, C.byref(ms)))
        return ms.value

    def mark_stop(self):
        check(self.dev.lib.mbx_timer_record_stop(self.dev.ctx, self.t))

    def elapsed(self):
        ms = C.c_float()
        check(self.dev.lib.mbx_timer_elapsed(self.t, C.byref(ms)))
        return ms.value

    def close(self):
        self.dev.lib.mbx_timer_destroy(self.t)


def bitmap_words(n):
    return (n + 31) // 32


def stride_words(n):
    return ((bitmap_words(n) + 63) // 64) * 64
