"""Minimal ctypes access to the HIP runtime that libgpp_hip.so itself loads
(libamdhip64.so.7), for tests that hand the library raw device pointers."""
import ctypes

import numpy as np

_rt = None
H2D, D2H = 1, 2


def rt():
    global _rt
    if _rt is None:
        import pygpp
        pygpp.load()  # resolve the same libamdhip64 instance the library uses
        lib = ctypes.CDLL("libamdhip64.so.7")
        lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        lib.hipFree.argtypes = [ctypes.c_void_p]
        lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        lib.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        lib.hipDeviceSynchronize.argtypes = []
        lib.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        lib.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        lib.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        lib.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int,
                                       ctypes.c_void_p]
        lib.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
        lib.hipHostFree.argtypes = [ctypes.c_void_p]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset", "hipDeviceSynchronize", "hipStreamCreate",
                  "hipStreamDestroy", "hipStreamSynchronize", "hipMemcpyAsync", "hipHostMalloc", "hipHostFree"):
            getattr(lib, f).restype = ctypes.c_int
        _rt = lib
    return _rt


class DeviceArray:
    """A device copy of a numpy array (freed on close / garbage collection)."""

    def __init__(self, host: np.ndarray):
        self.host = np.ascontiguousarray(host)
        self.nbytes = self.host.nbytes
        p = ctypes.c_void_p()
        assert rt().hipMalloc(ctypes.byref(p), max(self.nbytes, 16)) == 0
        self.ptr = p.value
        if self.nbytes:
            assert rt().hipMemcpy(self.ptr, self.host.ctypes.data, self.nbytes, H2D) == 0

    def get(self) -> np.ndarray:
        out = np.empty_like(self.host)
        assert rt().hipDeviceSynchronize() == 0
        assert rt().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H) == 0
        return out

    def close(self):
        if self.ptr:
            rt().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Stream:
    """A HIP stream of the current device (default priority)."""

    def __init__(self):
        h = ctypes.c_void_p()
        assert rt().hipStreamCreate(ctypes.byref(h)) == 0
        self.handle = h.value

    def synchronize(self):
        assert rt().hipStreamSynchronize(self.handle) == 0

    def close(self):
        if self.handle:
            rt().hipStreamDestroy(self.handle)
            self.handle = None


class PinnedArray:
    """Page-locked host memory (hipHostMalloc) viewed as a numpy array, for
    copies that the stream runs asynchronously."""

    def __init__(self, src: np.ndarray):
        src = np.ascontiguousarray(src)
        p = ctypes.c_void_p()
        assert rt().hipHostMalloc(ctypes.byref(p), max(src.nbytes, 16), 0) == 0
        self.ptr = p.value
        buf = (ctypes.c_char * max(src.nbytes, 16)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=src.dtype, count=src.size).reshape(src.shape)
        self.array[...] = src
        self.nbytes = src.nbytes

    def copy_to_device_async(self, dev_ptr: int, stream: Stream):
        assert rt().hipMemcpyAsync(dev_ptr, self.ptr, self.nbytes, H2D, stream.handle) == 0

    def close(self):
        if self.ptr:
            self.array = None
            rt().hipHostFree(self.ptr)
            self.ptr = None
