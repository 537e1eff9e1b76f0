"""Device buffers and events through the HIP runtime that libempanada_hip.so itself links
(/opt/rocm/lib/libamdhip64.so.7). torch ships its own HIP/HSA runtime; two runtimes in one process
cannot both open the GPU, so processes that load the library use this module instead of torch."""
import ctypes
import os

import numpy as np

_lib = None


def lib():
    global _lib
    if _lib is None:
        # the library's own dependency, so both share one runtime instance
        from . import native
        native.lib()
        for cand in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
            try:
                L = ctypes.CDLL(cand)
                break
            except OSError:
                continue
        else:
            raise RuntimeError("libamdhip64 not found")
        vp = ctypes.c_void_p
        L.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        L.hipFree.argtypes = [vp]
        L.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipStreamSynchronize.argtypes = [vp]
        L.hipEventCreate.argtypes = [ctypes.POINTER(vp)]
        L.hipEventDestroy.argtypes = [vp]
        L.hipEventRecord.argtypes = [vp, vp]
        L.hipEventSynchronize.argtypes = [vp]
        L.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), vp, vp]
        L.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.hipHostMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t, ctypes.c_uint]
        L.hipHostFree.argtypes = [vp]
        L.hipGetErrorString.restype = ctypes.c_char_p
        L.hipGetErrorString.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _check(rc, what):
    if rc:
        raise RuntimeError("%s failed: %s" % (what, lib().hipGetErrorString(rc).decode()))


H2D, D2H = 1, 2


def device_count():
    n = ctypes.c_int()
    rc = lib().hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.ptr = ctypes.c_void_p()
        _check(lib().hipMalloc(ctypes.byref(self.ptr), max(nbytes, 1)), "hipMalloc")

    @classmethod
    def from_array(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _check(lib().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return b

    def to_array(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        _check(lib().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    @property
    def address(self):
        return self.ptr.value

    def copy_from(self, a: np.ndarray, offset: int = 0):
        """Copies host array a to byte `offset` of the buffer."""
        a = np.ascontiguousarray(a)
        if offset < 0 or offset + a.nbytes > self.nbytes:
            raise ValueError("copy outside the buffer")
        _check(lib().hipMemcpy(ctypes.c_void_p(self.ptr.value + offset), a.ctypes.data, a.nbytes, H2D),
               "hipMemcpy H2D")

    def free(self):
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class HostBuffer:
    """Pinned host memory (hipHostMalloc), with a numpy view."""

    def __init__(self, nbytes):
        self.nbytes = nbytes
        self.ptr = ctypes.c_void_p()
        _check(lib().hipHostMalloc(ctypes.byref(self.ptr), max(nbytes, 1), 0), "hipHostMalloc")

    def array(self, dtype=np.uint8, shape=None):
        n = self.nbytes // np.dtype(dtype).itemsize
        a = np.ctypeslib.as_array(ctypes.cast(self.ptr, ctypes.POINTER(ctypes.c_uint8)), shape=(self.nbytes,))
        a = a.view(dtype)[:n]
        return a.reshape(shape) if shape is not None else a

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            lib().hipHostFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.h = ctypes.c_void_p()
        _check(lib().hipEventCreate(ctypes.byref(self.h)), "hipEventCreate")

    def record(self, stream=None):
        _check(lib().hipEventRecord(self.h, stream), "hipEventRecord")

    def synchronize(self):
        _check(lib().hipEventSynchronize(self.h), "hipEventSynchronize")

    def elapsed_ms(self, later: "Event") -> float:
        ms = ctypes.c_float()
        _check(lib().hipEventElapsedTime(ctypes.byref(ms), self.h, later.h), "hipEventElapsedTime")
        return ms.value


def copy_d2d(dst: DeviceBuffer, src: DeviceBuffer, nbytes: int):
    _check(lib().hipMemcpy(dst.ptr, src.ptr, nbytes, 3), "hipMemcpy D2D")


def synchronize():
    _check(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceJobs:
    """Read sets resident in HBM: one buffer, each set at a 256-byte aligned offset. jobs() gives
    (address, n, read length, min overlap) per set."""

    def __init__(self, sets):
        offs, total = [], 0
        for r, _ in sets:
            offs.append(total)
            total += (r.nbytes + 255) & ~255
        self.buf = DeviceBuffer(total)
        for (r, _), off in zip(sets, offs):
            if r.nbytes:
                self.buf.copy_from(r, off)
        self._jobs = [(self.buf.address + off, r.shape[0], r.shape[1] if r.ndim == 2 else 0, o)
                      for (r, o), off in zip(sets, offs)]

    def jobs(self):
        return self._jobs

    def free(self):
        self.buf.free()
