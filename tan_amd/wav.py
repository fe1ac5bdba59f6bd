"""WAV files for impulse responses and test signals (tan/common/wav.h),
read and written by the library (tanx_wav_*).

ReadWaveFile returns planar float samples scaled as the reference's reader
scales them (8-bit (u - 127) / 256, 16-bit s / 32768, 32-bit the stored
float); WriteWaveFileF clamps to [-1, 1] and converts toward zero after
scaling by 127 or 32767 (wav.cpp:240-318); WriteWaveFileS writes
interleaved 16-bit samples (wav.cpp:320-401).
"""
import ctypes
import os

import numpy as np

from . import _lib

AMF_OK = 0


class WavError(RuntimeError):
    def __init__(self, what, rc):
        super().__init__(f"{what}: AMF_RESULT {rc}")
        self.rc = rc


def _path(p):
    return os.fsencode(p)


def _info_dict(info):
    return {"sample_rate": info.sample_rate, "bits_per_sample": info.bits_per_sample,
            "channels": info.channels, "samples": info.samples, "format_tag": info.format_tag}


def ReadWaveInfo(path):
    """Header of a WAV file: sample_rate, bits_per_sample, channels, samples
    (per channel), format_tag."""
    lib = _lib.load()
    info = _lib.WavInfo()
    rc = lib.tanx_wav_read_info(_path(path), ctypes.byref(info))
    if rc != AMF_OK:
        raise WavError(f"ReadWaveInfo({path})", rc)
    return _info_dict(info)


def ReadWaveFile(path):
    """(info, samples): samples is a float32 array (channels, samples)."""
    lib = _lib.load()
    hdr = ReadWaveInfo(path)
    out = np.zeros((hdr["channels"], hdr["samples"]), np.float32)
    rows = (_lib._f32p * hdr["channels"])(
        *[out[c].ctypes.data_as(_lib._f32p) for c in range(hdr["channels"])])
    info = _lib.WavInfo()
    rc = lib.tanx_wav_read(_path(path), ctypes.byref(info), rows, out.shape[1])
    if rc != AMF_OK:
        raise WavError(f"ReadWaveFile({path})", rc)
    return _info_dict(info), out


def ReadWaveRaw(path):
    """(info, bytes): the interleaved sample bytes of the data chunk."""
    lib = _lib.load()
    hdr = ReadWaveInfo(path)
    n = hdr["samples"] * hdr["channels"] * (hdr["bits_per_sample"] // 8)
    buf = ctypes.create_string_buffer(max(1, n))
    info = _lib.WavInfo()
    rc = lib.tanx_wav_read_raw(_path(path), ctypes.byref(info), buf, n)
    if rc != AMF_OK:
        raise WavError(f"ReadWaveRaw({path})", rc)
    return _info_dict(info), buf.raw[:n]


def WriteWaveFileF(path, sample_rate, bits_per_sample, samples):
    """samples: (channels, n) float array (converted to float32)."""
    lib = _lib.load()
    x = np.ascontiguousarray(np.atleast_2d(samples), dtype=np.float32)
    rows = (_lib._f32p * x.shape[0])(*[x[c].ctypes.data_as(_lib._f32p) for c in range(x.shape[0])])
    rc = lib.tanx_wav_write(_path(path), sample_rate, x.shape[0], bits_per_sample, x.shape[1], rows)
    if rc != AMF_OK:
        raise WavError(f"WriteWaveFileF({path})", rc)


def WriteWaveFileS(path, sample_rate, channels, interleaved, bits_per_sample=16):
    """interleaved: int16 samples, frame by frame."""
    lib = _lib.load()
    s = np.ascontiguousarray(interleaved, dtype=np.int16)
    if s.size % channels:
        raise ValueError("sample count is not a multiple of the channel count")
    rc = lib.tanx_wav_write_short(_path(path), sample_rate, channels, bits_per_sample, s.size // channels,
                                  s.ctypes.data)
    if rc != AMF_OK:
        raise WavError(f"WriteWaveFileS({path})", rc)
