"""Minimal ctypes access to the HIP runtime (device buffers, copies, events)
for the benchmark and tests; the library itself manages its own memory."""
import ctypes
import weakref

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
            try:
                _hip = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipStreamDestroy.argtypes = [ctypes.c_void_p]
        _hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        _hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        _hip.hipMemcpyAsync.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]
        _hip.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t, ctypes.c_uint]
        _hip.hipHostFree.argtypes = [ctypes.c_void_p]
    return _hip


def check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed with HIP error %d" % (what, rc))


class DeviceArray:
    """A device allocation holding a copy of a numpy array."""

    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        check(hip().hipMalloc(ctypes.byref(self.ptr), max(1, nbytes)), "hipMalloc")
        self.nbytes = nbytes
        # freed when collected or by free(); not at interpreter exit, where a
        # queued update may still read it (the process exit reclaims it)
        self._fin = weakref.finalize(self, hip().hipFree, ctypes.c_void_p(self.ptr.value))
        self._fin.atexit = False

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        check(hip().hipMemcpy(d.ptr, ctypes.c_void_p(a.ctypes.data), a.nbytes, 1), "hipMemcpy H2D")
        return d

    def to_numpy(self, dtype, shape):
        out = np.empty(shape, dtype)
        check(hip().hipMemcpy(ctypes.c_void_p(out.ctypes.data), self.ptr, out.nbytes, 2), "hipMemcpy D2H")
        return out

    @property
    def value(self):
        return self.ptr.value

    def free(self):
        self._fin()
        self.ptr = ctypes.c_void_p()


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class _PinnedBlock:
    """Owner of a hipHostMalloc allocation (freed with the last numpy view)."""

    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        check(hip().hipHostMalloc(ctypes.byref(self.ptr), max(1, nbytes), 0), "hipHostMalloc")
        self.nbytes = nbytes
        # as DeviceArray: not freed at interpreter exit, where a query server
        # still draining its requests may write to it
        self._fin = weakref.finalize(self, hip().hipHostFree, ctypes.c_void_p(self.ptr.value))
        self._fin.atexit = False


def pinned_empty(shape, dtype):
    """numpy array in pinned host memory (asynchronous DMA source/target); the
    allocation lives as long as the array."""
    dtype = np.dtype(dtype)
    count = int(np.prod(shape))
    block = _PinnedBlock(count * dtype.itemsize)
    buf = (ctypes.c_char * max(1, count * dtype.itemsize)).from_address(block.ptr.value)
    buf._owner = block
    return np.frombuffer(buf, dtype=dtype, count=count).reshape(shape)


def pinned_copy(a):
    a = np.ascontiguousarray(a)
    out = pinned_empty(a.shape, a.dtype)
    out[...] = a
    return out
