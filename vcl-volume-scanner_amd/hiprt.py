"""Minimal ctypes access to the HIP runtime that libvscan.so itself links
(/opt/rocm/lib/libamdhip64.so): device buffers, streams and events for
callers of the C ABI that need their own device memory or timing.

One process must use one HIP runtime: torch-ROCm bundles its own copy, so it
is not loaded next to libvscan.so.
"""
import ctypes as C
import os

_hip = None
H2D, D2H, D2D = 1, 2, 3
HIP_NOT_READY = 600  # hipErrorNotReady


def hip():
    global _hip
    if _hip is None:
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")
        L = C.CDLL(path, mode=C.RTLD_GLOBAL)
        for f in ["hipMalloc", "hipFree", "hipMemcpy", "hipMemset", "hipDeviceSynchronize", "hipStreamCreate",
                  "hipStreamDestroy", "hipStreamSynchronize", "hipEventCreate", "hipEventDestroy", "hipEventRecord",
                  "hipEventSynchronize", "hipEventElapsedTime", "hipGetDeviceCount", "hipMemGetInfo", "hipEventQuery",
                  "hipHostMalloc", "hipHostFree"]:
            getattr(L, f).restype = C.c_int
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
        L.hipHostFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipStreamDestroy.argtypes = [C.c_void_p]
        L.hipStreamSynchronize.argtypes = [C.c_void_p]
        L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipEventDestroy.argtypes = [C.c_void_p]
        L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        L.hipEventSynchronize.argtypes = [C.c_void_p]
        L.hipEventQuery.argtypes = [C.c_void_p]
        L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        L.hipGetDeviceCount.argtypes = [C.POINTER(C.c_int)]
        L.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        _hip = L
    return _hip


def _ok(code, what):
    if code != 0:
        raise RuntimeError(f"{what} failed with hipError {code}")


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(self.ptr), self.nbytes), "hipMalloc")

    def upload(self, arr):
        assert arr.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(self.ptr, arr.ctypes.data_as(C.c_void_p), arr.nbytes, H2D), "hipMemcpy H2D")

    def download(self, arr):
        assert arr.nbytes <= self.nbytes
        _ok(hip().hipMemcpy(arr.ctypes.data_as(C.c_void_p), self.ptr, arr.nbytes, D2H), "hipMemcpy D2H")
        return arr

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()


class PinnedBuffer:
    """Page-locked host memory (hipHostMalloc) with a numpy view: frames
    handed to the library from here upload by DMA alone, with no staging
    copy on the CPU."""

    def __init__(self, shape, dtype):
        import numpy as np
        self.nbytes = int(np.prod(shape)) * np.dtype(dtype).itemsize
        self.ptr = C.c_void_p()
        _ok(hip().hipHostMalloc(C.byref(self.ptr), max(1, self.nbytes), 0), "hipHostMalloc")
        raw = (C.c_uint8 * self.nbytes).from_address(self.ptr.value)
        self.array = np.frombuffer(raw, dtype=dtype).reshape(shape)

    def free(self):
        if self.ptr:
            self.array = None
            hip().hipHostFree(self.ptr)
            self.ptr = C.c_void_p()


class Stream:
    def __init__(self):
        self.h = C.c_void_p()
        _ok(hip().hipStreamCreate(C.byref(self.h)), "hipStreamCreate")

    def sync(self):
        _ok(hip().hipStreamSynchronize(self.h), "hipStreamSynchronize")

    def destroy(self):
        if self.h:
            hip().hipStreamDestroy(self.h)
            self.h = C.c_void_p()


class Event:
    def __init__(self):
        self.h = C.c_void_p()
        _ok(hip().hipEventCreate(C.byref(self.h)), "hipEventCreate")

    def record(self, stream=None):
        _ok(hip().hipEventRecord(self.h, stream.h if stream is not None else None), "hipEventRecord")

    def sync(self):
        _ok(hip().hipEventSynchronize(self.h), "hipEventSynchronize")

    def query(self):
        """True once the work before the event has completed."""
        r = hip().hipEventQuery(self.h)
        if r not in (0, HIP_NOT_READY):
            _ok(r, "hipEventQuery")
        return r == 0

    def elapsed_ms(self, later):
        ms = C.c_float()
        _ok(hip().hipEventElapsedTime(C.byref(ms), self.h, later.h), "hipEventElapsedTime")
        return ms.value

    def destroy(self):
        if self.h:
            hip().hipEventDestroy(self.h)
            self.h = C.c_void_p()


def copy_d2d(dst, src, nbytes):
    """Device-to-device copy between two DeviceBuffers (synchronous)."""
    assert nbytes <= dst.nbytes and nbytes <= src.nbytes
    _ok(hip().hipMemcpy(dst.ptr, src.ptr, nbytes, D2D), "hipMemcpy D2D")


def device_sync():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    free, total = C.c_size_t(), C.c_size_t()
    _ok(hip().hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value, total.value
