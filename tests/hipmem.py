"""Device buffers for GPU tests through the HIP runtime that libmpt.so itself
links (ctypes on libamdhip64.so): allocate, copy in and out, free. Test
plumbing only; the library under test never sees anything but pointers."""
import ctypes as C

import numpy as np

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        from cadise_amd import native
        native.load_library()
        # by soname: the HIP runtime already loaded for libmpt.so (loading another
        # copy of the runtime would see no device)
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipDeviceSynchronize.argtypes = []
        L.hipSetDevice.argtypes = [C.c_int]
        _hip = L
    return _hip


class DeviceBuffer:
    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        rc = hip().hipMalloc(C.byref(self.ptr), max(16, self.nbytes))
        if rc != 0:
            raise MemoryError(f"hipMalloc({self.nbytes}) failed: {rc}")

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def download(self, a: np.ndarray) -> np.ndarray:
        assert a.flags.c_contiguous and a.nbytes <= self.nbytes
        assert hip().hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, D2H) == 0
        return a

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        self.free()


def synchronize():
    assert hip().hipDeviceSynchronize() == 0
