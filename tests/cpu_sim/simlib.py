"""The CPU simulation build of the library (tests/cpu_sim/build/libtan_sim.so)
bound with the same signatures as the HIP library, and guarded "device"
buffers for its device-pointer entry points.

Test infrastructure: the simulation runs the library's own kernel and engine
source on the host, one workgroup at a time, with every device allocation
fenced by guard pages (see sim_runtime.cpp).  It checks memory safety and
synchronisation of launches on the CPU before they reach a GPU.
"""
import ctypes
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB_PATH = os.path.join(HERE, "build", "libtan_sim.so")
sys.path.insert(0, ROOT)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def load():
    """The simulation library, typed like tan_amd._lib, and installed as the
    library tan_amd's classes use in this process."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    from tan_amd import _lib as L
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in L._SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    lib.sim_device_alloc.restype = ctypes.c_void_p
    lib.sim_device_alloc.argtypes = [ctypes.c_size_t]
    lib.sim_device_free.argtypes = [ctypes.c_void_p]
    lib.sim_live_allocations.restype = ctypes.c_size_t
    L._lib = lib
    _lib = lib
    return lib


class Dev:
    """A guarded device buffer of float32 (rows, cols) with a numpy view."""

    def __init__(self, rows, cols, fill=None):
        lib = load()
        self.shape = (rows, cols)
        self.ptr = lib.sim_device_alloc(4 * rows * cols)
        if not self.ptr:
            raise MemoryError("sim_device_alloc failed")
        buf = (ctypes.c_float * (rows * cols)).from_address(self.ptr)
        self.a = np.frombuffer(buf, dtype=np.float32).reshape(rows, cols)
        if fill is not None:
            self.a[...] = fill

    def at(self, col):
        """Address of column `col` of row 0."""
        return self.ptr + 4 * col

    @property
    def stride(self):
        return self.shape[1]

    def free(self):
        if self.ptr:
            load().sim_device_free(self.ptr)
            self.ptr = None
