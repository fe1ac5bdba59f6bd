"""ctypes view of the CPU oracle (oracle/tan_oracle.c).

Test infrastructure: the oracle is the checker, never the thing measured.
"""
import ctypes
import os

import numpy as np

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_LIB_PATH = os.path.join(_ROOT, "oracle", "build", "libtan_oracle.so")

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u32p = ctypes.POINTER(ctypes.c_uint32)

AMF_OK = 0
AMF_INVALID_ARG = 4
AMF_NOT_FOUND = 11
AMF_WRONG_STATE = 15

XFADE_INTENDED = 0
XFADE_REFERENCE = 1


def _load():
    if not os.path.exists(_LIB_PATH):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", os.path.join(_ROOT, "oracle")])
    lib = ctypes.CDLL(_LIB_PATH)
    lib.orc_dft_f64.argtypes = [ctypes.c_int, _f64p, _f64p, ctypes.c_int]
    lib.orc_direct_conv_f64.argtypes = [_f32p, ctypes.c_long, _f32p, ctypes.c_long,
                                        _f64p, ctypes.c_long]
    lib.orc_fft_transform.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(_f32p), ctypes.POINTER(_f32p)]
    lib.orc_conv_create.restype = ctypes.c_void_p
    lib.orc_conv_create.argtypes = [ctypes.c_int] * 5
    lib.orc_conv_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_conv_update_td.argtypes = [ctypes.c_void_p, ctypes.POINTER(_f32p), ctypes.c_size_t,
                                       _u32p, ctypes.c_uint32]
    lib.orc_conv_process.argtypes = [ctypes.c_void_p, ctypes.POINTER(_f32p),
                                     ctypes.POINTER(_f32p), ctypes.c_size_t, _u32p,
                                     ctypes.POINTER(ctypes.c_size_t)]
    lib.orc_conv_finalize.argtypes = [ctypes.c_void_p]
    lib.orc_conv_get_next_free_channel.argtypes = [ctypes.c_void_p, _u32p, _u32p]
    lib.orc_conv_last_was_crossfade.argtypes = [ctypes.c_void_p]
    lib.orc_set_mutation.argtypes = [ctypes.c_int]
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    lib.orc_num_procs.argtypes = []
    lib.orc_mix.argtypes = [ctypes.POINTER(_f32p), ctypes.c_int, ctypes.c_size_t, _f32p]
    lib.orc_convert.restype = ctypes.c_int
    lib.orc_convert.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t, ctypes.c_float,
                                ctypes.POINTER(ctypes.c_int)]
    lib.orc_room_ear_params.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_uint, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.POINTER(ctypes.c_longlong)]
    lib.orc_room_image.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
    lib.orc_room_response.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_uint, ctypes.c_int, _f32p, _f32p]
    lib.orc_room_head_model.argtypes = [ctypes.c_float, ctypes.c_void_p]
    lib.orc_bench_upols.restype = ctypes.c_double
    lib.orc_bench_upols.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_uint64, _f64p]
    return lib


lib = _load()


def _rows(a):
    """Array of row pointers for a C-contiguous 2-D float32 array."""
    assert a.dtype == np.float32 and a.flags.c_contiguous and a.ndim == 2
    base = a.ctypes.data
    stride = a.strides[0]
    return (_f32p * a.shape[0])(*[ctypes.cast(base + i * stride, _f32p) for i in range(a.shape[0])])


def _flags(flags, C):
    if flags is None:
        return None
    arr = np.ascontiguousarray(flags, dtype=np.uint32)
    assert arr.shape == (C,)
    return arr


def dft_f64(x, inverse=False):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib.orc_dft_f64(x.size, x.ctypes.data_as(_f64p), out.ctypes.data_as(_f64p), int(inverse))
    return out


def direct_conv_f64(x, h, ny=None):
    x = np.ascontiguousarray(x, dtype=np.float32)
    h = np.ascontiguousarray(h, dtype=np.float32)
    ny = x.size if ny is None else ny
    y = np.empty(ny, dtype=np.float64)
    lib.orc_direct_conv_f64(x.ctypes.data_as(_f32p), x.size, h.ctypes.data_as(_f32p), h.size,
                            y.ctypes.data_as(_f64p), ny)
    return y


def fft_transform(direction, log2len, inp):
    """TANFFT::Transform semantics; inp is (channels, 2*2^log2len) float32."""
    inp = np.ascontiguousarray(inp, dtype=np.float32)
    out = np.zeros_like(inp)
    rc = lib.orc_fft_transform(direction, log2len, inp.shape[0], _rows(inp), _rows(out))
    return rc, out


class OracleConvolution:
    """The restated CPU TANConvolution (FFT_PARTITIONED_UNIFORM)."""

    def __init__(self, L, B, C, xfade=XFADE_INTENDED, finalize_mode=False):
        self.L, self.B, self.C = L, B, C
        self.h = lib.orc_conv_create(L, B, C, xfade, int(finalize_mode))
        if not self.h:
            raise ValueError("bad oracle convolution arguments")

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_conv_destroy(self.h)
            self.h = None

    def update_td(self, ir, n=None, flags=None, opflags=0):
        ir = np.ascontiguousarray(ir, dtype=np.float32)
        n = ir.shape[1] if n is None else n
        f = _flags(flags, self.C)
        return lib.orc_conv_update_td(self.h, _rows(ir), n,
                                      None if f is None else f.ctypes.data_as(_u32p), opflags)

    def process(self, x, out=None, flags=None, n=None):
        x = np.ascontiguousarray(x, dtype=np.float32)
        if out is None:
            out = np.zeros((self.C, self.B), dtype=np.float32)
        n = x.shape[1] if n is None else n
        done = ctypes.c_size_t(0)
        f = _flags(flags, self.C)
        rc = lib.orc_conv_process(self.h, _rows(x), _rows(out), n,
                                  None if f is None else f.ctypes.data_as(_u32p),
                                  ctypes.byref(done))
        return rc, out, done.value

    def finalize(self):
        return lib.orc_conv_finalize(self.h)

    def last_was_crossfade(self):
        return bool(lib.orc_conv_last_was_crossfade(self.h))

    def next_free_channel(self, flags):
        idx = ctypes.c_uint32(0)
        f = _flags(flags, self.C)
        rc = lib.orc_conv_get_next_free_channel(self.h, ctypes.byref(idx),
                                                None if f is None else f.ctypes.data_as(_u32p))
        return rc, idx.value


def set_mutation(m):
    lib.orc_set_mutation(m)


def set_threads(n):
    return lib.orc_set_threads(n)


def mix(rows, n=None):
    """orc_mix over the rows of a (C, >= n) float32 array."""
    a = np.ascontiguousarray(rows, dtype=np.float32)
    n = a.shape[1] if n is None else n
    out = np.zeros(n, np.float32)
    lib.orc_mix(_rows(a), a.shape[0], n, out.ctypes.data_as(_f32p))
    return out


def convert(inp, in_step, in_offset, out, out_step, out_offset, n, gain):
    """orc_convert between numpy float32 / int16 arrays (out is written in
    place); returns the clip flag."""
    t = {np.dtype(np.float32): 0, np.dtype(np.int16): 1}
    clipped = ctypes.c_int(0)
    rc = lib.orc_convert(inp.ctypes.data, in_step, in_offset, t[inp.dtype], out.ctypes.data, out_step, out_offset,
                         t[out.dtype], n, gain, ctypes.byref(clipped))
    assert rc == 0
    return bool(clipped.value)


# ---- 5. room responses (tan_oracle_room.c) ----

class RoomEar(ctypes.Structure):
    """orc_room_ear: the kernel arguments of one ear."""
    _fields_ = [("src", ctypes.c_float * 3), ("head", ctypes.c_float * 3), ("ev", ctypes.c_float * 3),
                ("ev_norm", ctypes.c_float), ("dim", ctypes.c_float * 3), ("damp", ctypes.c_float * 6),
                ("max_gain", ctypes.c_float), ("d_min", ctypes.c_float), ("fs", ctypes.c_int),
                ("len", ctypes.c_int), ("hrtf_len", ctypes.c_int), ("filt_len", ctypes.c_int),
                ("n", ctypes.c_int * 3), ("suppress_direct", ctypes.c_int),
                ("lp", ctypes.c_float * 64), ("hp", ctypes.c_float * 64)]


def room_head_model(ear_spacing, listener):
    """Fills the listener struct's head filters (any struct with the
    StereoListener layout)."""
    return lib.orc_room_head_model(ear_spacing, ctypes.byref(listener))


def room_ear_params(room, source, listener, fs, length, flags=0, max_bounces=0):
    ears = (RoomEar * 2)()
    images = ctypes.c_longlong(0)
    rc = lib.orc_room_ear_params(ctypes.byref(room), ctypes.byref(source), ctypes.byref(listener), fs, length,
                                 flags, max_bounces, ears, ctypes.byref(images))
    return rc, ears, images.value


def room_image(ear, ix, iy, iz):
    """(kind, index, taps, scale, hf): kind 0 none, 1 one tap, 2 head-filtered."""
    idx, taps = ctypes.c_int(), ctypes.c_int()
    sc, hf = ctypes.c_float(), ctypes.c_float()
    kind = lib.orc_room_image(ctypes.byref(ear), ix, iy, iz, ctypes.byref(idx), ctypes.byref(taps),
                              ctypes.byref(sc), ctypes.byref(hf))
    return kind, idx.value, taps.value, sc.value, hf.value


def room_response(room, source, listener, fs, length, flags=0, max_bounces=0):
    """(rc, left, right) float32."""
    left = np.zeros(length, np.float32)
    right = np.zeros(length, np.float32)
    rc = lib.orc_room_response(ctypes.byref(room), ctypes.byref(source), ctypes.byref(listener), fs, length,
                               flags, max_bounces, left.ctypes.data_as(_f32p), right.ctypes.data_as(_f32p))
    return rc, left, right
