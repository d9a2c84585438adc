"""Minimal ctypes binding of the HIP runtime (libamdhip64): device buffers,
streams and events, for callers of the C ABI that hold raw device pointers
(the drop-in entry points take plain float*)."""
import ctypes as C

import numpy as np

_hip = None

H2D, D2H, D2D = 1, 2, 3


class HipError(RuntimeError):
    pass


def rt():
    global _hip
    if _hip is None:
        L = C.CDLL("libamdhip64.so")
        vp = C.c_void_p
        L.hipMalloc.argtypes = [C.POINTER(vp), C.c_size_t]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
        L.hipMemcpyAsync.argtypes = [vp, vp, C.c_size_t, C.c_int, vp]
        L.hipMemset.argtypes = [vp, C.c_int, C.c_size_t]
        L.hipDeviceSynchronize.argtypes = []
        L.hipGetErrorString.argtypes = [C.c_int]
        L.hipGetErrorString.restype = C.c_char_p
        L.hipStreamCreate.argtypes = [C.POINTER(vp)]
        L.hipStreamDestroy.argtypes = [vp]
        L.hipStreamSynchronize.argtypes = [vp]
        L.hipEventCreate.argtypes = [C.POINTER(vp)]
        L.hipEventDestroy.argtypes = [vp]
        L.hipEventRecord.argtypes = [vp, vp]
        L.hipEventSynchronize.argtypes = [vp]
        L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
        L.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        L.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        _hip = L
    return _hip


def ok(rc, what):
    if rc != 0:
        raise HipError(f"{what}: {rt().hipGetErrorString(rc).decode()} ({rc})")


def device_count():
    n = C.c_int(0)
    rc = rt().hipGetDeviceCount(C.byref(n))
    return n.value if rc == 0 else 0


def sync():
    ok(rt().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    f, t = C.c_size_t(), C.c_size_t()
    ok(rt().hipMemGetInfo(C.byref(f), C.byref(t)), "hipMemGetInfo")
    return f.value, t.value


class DeviceArray:
    """A device buffer of float32 (or another numpy dtype)."""

    def __init__(self, n, dtype=np.float32, zero=True):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        ok(rt().hipMalloc(C.byref(p), max(self.nbytes, 4)), "hipMalloc")
        self.ptr = p.value
        if zero:
            ok(rt().hipMemset(self.ptr, 0, max(self.nbytes, 4)), "hipMemset")

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype, zero=False)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.n
        ok(rt().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, H2D), "hipMemcpy H2D")

    def download(self):
        out = np.empty(self.n, self.dtype)
        ok(rt().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        ok(rt().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def sync(self):
        ok(rt().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            rt().hipStreamDestroy(self.handle)
            self.handle = None


class Event:
    def __init__(self):
        e = C.c_void_p()
        ok(rt().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None):
        ok(rt().hipEventRecord(self.handle, stream.handle if isinstance(stream, Stream) else stream), "hipEventRecord")

    def sync(self):
        ok(rt().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, later):
        t = C.c_float()
        ok(rt().hipEventElapsedTime(C.byref(t), self.handle, later.handle), "hipEventElapsedTime")
        return t.value
