"""Loader for the in-tree HIP library (tan_amd/lib/libtan_mi355x.so).

The library is the product: there is no CPU fallback.  A missing or
unloadable library raises immediately.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libtan_mi355x.so")

_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_u32 = ctypes.c_uint32
_sz = ctypes.c_size_t


class ConvInfo(ctypes.Structure):
    _fields_ = [("ir_len", ctypes.c_uint32), ("block", ctypes.c_uint32),
                ("channels", ctypes.c_uint32), ("partitions", ctypes.c_uint32),
                ("tail_chunks", ctypes.c_uint32), ("finalize_mode", ctypes.c_uint32),
                ("device_bytes", ctypes.c_uint64), ("ticks", ctypes.c_uint64),
                ("crossfades", ctypes.c_uint64), ("ir_precision", ctypes.c_uint32),
                ("nu_multiple", ctypes.c_uint32), ("pass_blocks", ctypes.c_uint32),
                ("cached_rows", ctypes.c_uint32), ("ir_slots", ctypes.c_uint32)]


class ConvOptions(ctypes.Structure):
    _fields_ = [("ir_precision", ctypes.c_uint32), ("nu_multiple", ctypes.c_uint32),
                ("nu_phase", ctypes.c_uint32), ("pass_blocks", ctypes.c_uint32),
                ("pass_phase", ctypes.c_uint32), ("max_tail_chunks", ctypes.c_uint32),
                ("ir_slots", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 1)]


class WavInfo(ctypes.Structure):
    _fields_ = [("sample_rate", ctypes.c_uint32), ("bits_per_sample", ctypes.c_uint16),
                ("channels", ctypes.c_uint16), ("samples", ctypes.c_uint32),
                ("format_tag", ctypes.c_uint16), ("reserved", ctypes.c_uint16)]


class KernelTiming(ctypes.Structure):
    _fields_ = [("head_ms", ctypes.c_double), ("tail_ms", ctypes.c_double),
                ("head_launches", ctypes.c_uint64), ("tail_launches", ctypes.c_uint64),
                ("head_alg_bytes", ctypes.c_uint64), ("tail_alg_bytes", ctypes.c_uint64),
                ("blocks", ctypes.c_uint64)]


_SIGS = {
    "tanx_ctx_create": (_i32, [ctypes.c_uint64, _i32, ctypes.POINTER(_vp)]),
    "tanx_ctx_create_on_streams": (_i32, [ctypes.c_uint64, _i32, _vp, _vp, ctypes.POINTER(_vp)]),
    "tanx_ctx_destroy": (_i32, [_vp]),
    "tanx_ctx_device": (_i32, [_vp, ctypes.POINTER(_i32)]),
    "tanx_ctx_process_stream": (_vp, [_vp]),
    "tanx_ctx_update_stream": (_vp, [_vp]),
    "tanx_mixer_create": (_i32, [_vp, ctypes.c_size_t, _i32, ctypes.POINTER(_vp)]),
    "tanx_mixer_destroy": (_i32, [_vp]),
    "tanx_mix": (_i32, [_vp, _vp, _vp]),
    "tanx_mix_device": (_i32, [_vp, _vp, _vp]),
    "tanx_mix_device_strided": (_i32, [_vp, _vp, ctypes.c_size_t, _vp]),
    "tanx_convert_device": (_i32, [_vp, _vp, ctypes.c_size_t, ctypes.c_size_t, _i32, _vp, ctypes.c_size_t,
                                   ctypes.c_size_t, _i32, ctypes.c_size_t, ctypes.c_float, ctypes.POINTER(_i32)]),
    "tanx_convert": (_i32, [_vp, _vp, ctypes.c_size_t, _i32, _vp, ctypes.c_size_t, _i32, ctypes.c_size_t,
                            ctypes.c_float, ctypes.POINTER(_i32)]),
    "tanx_conv_create": (_i32, [_vp, _u32, _u32, _u32, _u32, ctypes.POINTER(_vp)]),
    "tanx_conv_create_ex": (_i32, [_vp, _u32, _u32, _u32, _u32, ctypes.c_void_p, ctypes.POINTER(_vp)]),
    "tanx_conv_destroy": (_i32, [_vp]),
    "tanx_conv_estimate_bytes": (_i32, [_u32, _u32, _u32, _u32, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
    "tanx_conv_update_td": (_i32, [_vp, ctypes.POINTER(_f32p), _sz, _u32p, _u32]),
    "tanx_conv_update_td_device": (_i32, [_vp, _vp, _sz, _sz, _u32p, _u32]),
    "tanx_conv_process": (_i32, [_vp, ctypes.POINTER(_f32p), ctypes.POINTER(_f32p), _sz, _u32p,
                                 ctypes.POINTER(_sz)]),
    "tanx_conv_process_device": (_i32, [_vp, _vp, _sz, _vp, _sz, _sz, _u32p, ctypes.POINTER(_sz)]),
    "tanx_conv_finalize": (_i32, [_vp]),
    "tanx_dft_r2c_mfma_device": (_i32, [_vp, _u32, _u32, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t, _u32]),
    "tanx_conv_process_direct": (_i32, [_vp, _vp, _vp, _vp, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t),
                                        ctypes.c_void_p]),
    "tanx_conv_process_direct_device": (_i32, [_vp, _vp, _sz, _vp, _sz, _vp, _sz, _sz, ctypes.POINTER(_sz),
                                               ctypes.c_void_p]),
    "tanx_group_update_td_device": (_i32, [_u32, _vp, _vp, _vp, _vp, _vp, _u32]),
    "tanx_conv_flush": (_i32, [_vp, _u32]),
    "tanx_conv_get_next_free_channel": (_i32, [_vp, _u32p, _u32p]),
    "tanx_conv_synchronize": (_i32, [_vp]),
    "tanx_conv_get_info": (_i32, [_vp, ctypes.POINTER(ConvInfo)]),
    "tanx_conv_set_tail_chunks": (_i32, [_vp, _u32]),
    "tanx_conv_last_was_crossfade": (_i32, [_vp]),
    "tanx_conv_enable_timing": (_i32, [_vp, _u32]),
    "tanx_conv_set_fused": (_i32, [_vp, _u32]),
    "tanx_conv_set_graph": (_i32, [_vp, _u32]),
    "tanx_conv_autotune": (_i32, [_vp, ctypes.POINTER(_u32), ctypes.POINTER(ctypes.c_double)]),
    "tanx_conv_read_timing": (_i32, [_vp, ctypes.POINTER(KernelTiming), _u32]),
    "tanx_conv_process_stream": (_vp, [_vp]),
    "tanx_ctx_enable_group_timing": (_i32, [_vp, _u32]),
    "tanx_ctx_set_group_queue": (_i32, [_vp, _u32]),
    "tanx_ctx_read_group_timing": (_i32, [_vp, ctypes.POINTER(KernelTiming), _u32]),
    "tanx_group_process_device": (_i32, [_u32, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                         ctypes.POINTER(_sz), ctypes.POINTER(_vp), ctypes.POINTER(_sz),
                                         ctypes.c_void_p]),
    "tanx_group_process_device_noalloc": (_i32, [_u32, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                 ctypes.POINTER(_sz), ctypes.POINTER(_vp), ctypes.POINTER(_sz),
                                                 ctypes.c_void_p]),
    "tanx_fft_transform": (_i32, [_vp, _u32, _u32, _u32, ctypes.POINTER(_f32p), ctypes.POINTER(_f32p)]),
    "tanx_fft_transform_device": (_i32, [_vp, _u32, _u32, _u32, _vp, _vp, _sz]),
    "tanx_version": (ctypes.c_char_p, []),
    "tanx_room_head_model": (_i32, [ctypes.c_float, _vp]),
    "tanx_room_response_device": (_i32, [_vp, _u32, _vp, _vp, _vp, _i32, _i32, _u32, _i32, _vp, _sz]),
    "tanx_room_response": (_i32, [_vp, _vp, _vp, _vp, _i32, _i32, _u32, _i32, _f32p, _f32p]),
    "tanx_wav_read_info": (_i32, [ctypes.c_char_p, ctypes.POINTER(WavInfo)]),
    "tanx_wav_read": (_i32, [ctypes.c_char_p, ctypes.POINTER(WavInfo), ctypes.POINTER(_f32p), _sz]),
    "tanx_wav_read_raw": (_i32, [ctypes.c_char_p, ctypes.POINTER(WavInfo), _vp, _sz]),
    "tanx_wav_write": (_i32, [ctypes.c_char_p, _u32, ctypes.c_uint16, ctypes.c_uint16, _u32,
                              ctypes.POINTER(_f32p)]),
    "tanx_wav_write_short": (_i32, [ctypes.c_char_p, _u32, ctypes.c_uint16, ctypes.c_uint16, _u32, _vp]),
}

_lib = None


def _bind_process_hip_runtime():
    """PyTorch-ROCm wheels carry their own libamdhip64/libhsa-runtime64 and
    request them by unversioned name; this library requests the versioned
    sonames (libamdhip64.so.7, libhsa-runtime64.so.1), which those copies
    also carry.  Importing torch first makes the dynamic loader bind this
    library to the runtime torch already loaded, so one HIP runtime serves
    the process; loading this library first would make torch load a second
    runtime that then finds no devices."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load():
    """Load and type the library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    _bind_process_hip_runtime()
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"tan_amd: HIP library not built ({LIB_PATH}); run __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return list(_SIGS)
