"""Minimal HIP runtime access through ctypes: device buffers, streams and events for
host code that hands device pointers to the C ABI (no torch needed)."""
import ctypes as C
import os

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so")):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipGetErrorString.restype = C.c_char_p
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipEventSynchronize.argtypes = [C.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventDestroy.argtypes = [C.c_void_p]
        _hip.hipStreamDestroy.argtypes = [C.c_void_p]
    return _hip


def check(rc, what=""):
    if rc != 0:
        raise RuntimeError("%s: %s" % (what, hip().hipGetErrorString(rc).decode()))


H2D, D2H = 1, 2


class DeviceBuffer:
    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        self.nbytes = nbytes
        check(hip().hipMalloc(C.byref(self.ptr), nbytes), "hipMalloc")

    def zero(self):
        check(hip().hipMemset(self.ptr, 0, self.nbytes), "hipMemset")

    def to_host(self, arr):
        check(hip().hipMemcpy(arr.ctypes.data_as(C.c_void_p), self.ptr, arr.nbytes, D2H), "hipMemcpy D2H")
        return arr

    def from_host(self, arr):
        check(hip().hipMemcpy(self.ptr, arr.ctypes.data_as(C.c_void_p), arr.nbytes, H2D), "hipMemcpy H2D")

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()


class Stream:
    def __init__(self):
        self.ptr = C.c_void_p()
        check(hip().hipStreamCreate(C.byref(self.ptr)), "hipStreamCreate")

    def synchronize(self):
        check(hip().hipStreamSynchronize(self.ptr), "hipStreamSynchronize")

    @property
    def address(self):
        return self.ptr.value


class Event:
    def __init__(self):
        self.ptr = C.c_void_p()
        check(hip().hipEventCreate(C.byref(self.ptr)), "hipEventCreate")

    def record(self, stream):
        check(hip().hipEventRecord(self.ptr, stream.ptr), "hipEventRecord")

    def synchronize(self):
        check(hip().hipEventSynchronize(self.ptr), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = C.c_float()
        check(hip().hipEventElapsedTime(C.byref(ms), self.ptr, end.ptr), "hipEventElapsedTime")
        return ms.value


D2D = 3


def copy_device(dst, src, nbytes):
    check(hip().hipMemcpy(dst.ptr, src.ptr, nbytes, D2D), "hipMemcpy D2D")


def device_synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
