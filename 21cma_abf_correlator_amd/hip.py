"""Minimal ctypes binding of the HIP runtime that libcorr_cuda.so links.

Device buffers, streams and events for the engine-level entry points, taken
from the very runtime the library uses. It is opened by soname after the
library, so the dynamic loader hands back the copy already mapped for it.
"""

import ctypes

import numpy as np

H2D, D2H, D2D = 1, 2, 3
_rt = None


def rt():
    global _rt
    if _rt is None:
        from . import corr
        corr.load()  # maps the library's own runtime first
        L = ctypes.CDLL("libamdhip64.so.7")
        vp, sz, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        for name, args in {
            "hipMalloc": [ctypes.POINTER(vp), sz],
            "hipFree": [vp],
            "hipMemcpy": [vp, vp, sz, i32],
            "hipMemcpyAsync": [vp, vp, sz, i32, vp],
            "hipMemset": [vp, i32, sz],
            "hipMemsetAsync": [vp, i32, sz, vp],
            "hipDeviceSynchronize": [],
            "hipGetDeviceCount": [ctypes.POINTER(i32)],
            "hipSetDevice": [i32],
            "hipStreamCreate": [ctypes.POINTER(vp)],
            "hipStreamDestroy": [vp],
            "hipStreamSynchronize": [vp],
            "hipEventCreate": [ctypes.POINTER(vp)],
            "hipEventDestroy": [vp],
            "hipEventRecord": [vp, vp],
            "hipEventSynchronize": [vp],
            "hipEventElapsedTime": [ctypes.POINTER(ctypes.c_float), vp, vp],
        }.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = i32
        L.hipGetErrorString.argtypes = [i32]
        L.hipGetErrorString.restype = ctypes.c_char_p
        _rt = L
    return _rt


def check(err, what):
    if err != 0:
        raise RuntimeError(f"{what}: HIP error {err}: {rt().hipGetErrorString(err).decode()}")


def device_count():
    n = ctypes.c_int(0)
    err = rt().hipGetDeviceCount(ctypes.byref(n))
    return n.value if err == 0 else 0


def synchronize():
    check(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Buffer:
    """A device allocation of `nbytes` bytes."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        check(rt().hipMalloc(ctypes.byref(p), max(1, int(nbytes))), f"hipMalloc({nbytes})")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a, offset=0):
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        check(rt().hipMemcpy(self.ptr + offset, a.ctypes.data, a.nbytes, H2D), "hipMemcpy(H2D)")

    def download(self, a, offset=0):
        assert a.flags["C_CONTIGUOUS"] and offset + a.nbytes <= self.nbytes
        check(rt().hipMemcpy(a.ctypes.data, self.ptr + offset, a.nbytes, D2H), "hipMemcpy(D2H)")
        return a

    def to_numpy(self, dtype, shape, offset=0):
        out = np.empty(shape, dtype=dtype)
        return self.download(out, offset)

    def zero(self):
        check(rt().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def free(self):
        if self.ptr is not None:
            rt().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def copy_d2d(dst, src, nbytes):
    check(rt().hipMemcpy(dst, src, nbytes, D2D), "hipMemcpy(D2D)")


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        check(rt().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        check(rt().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle is not None:
            rt().hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        check(rt().hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream):
        check(rt().hipEventRecord(self.handle, stream.handle if stream else None), "hipEventRecord")

    def synchronize(self):
        check(rt().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        check(rt().hipEventElapsedTime(ctypes.byref(ms), self.handle, later.handle),
              "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self.handle is not None:
                rt().hipEventDestroy(self.handle)
        except Exception:
            pass
