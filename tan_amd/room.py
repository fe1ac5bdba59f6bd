"""Room impulse responses on the GPU: TrueAudioVR's image-source generator
(samples/src/TrueAudioVR/TrueAudioVR.h:128-213) on the library's HIP
kernels (tanx_room_*), as an IR source for TANConvolution.

RoomDefinition, MonoSource and StereoListener keep the reference's struct
layouts; generateRoomResponse fills host arrays (numpy) or device tensors
(torch CUDA), and GenerateRoomResponses computes many responses in one
launch into one device tensor, ready for UpdateResponseTD on device memory.
"""
import ctypes

import numpy as np

from . import _lib

AMF_OK = 0
GENROOM_NONE = 0
GENROOM_LIMIT_BOUNCES = 0x40
GENROOM_SUPPRESS_DIRECT = 0x80
HEAD_FILTER_MAX = 64
SPEED_OF_SOUND = 340.0  # AmdTrueAudioVR::S, the speed the image arrival times use


class RoomDefinition(ctypes.Structure):
    """Dimensions in meters (x: width, y: height, z: length) and the wall
    damping factors in [0, 1] (TrueAudioVR.h:101-107)."""
    _fields_ = [("width", ctypes.c_float), ("height", ctypes.c_float), ("length", ctypes.c_float),
                ("damp_left", ctypes.c_float), ("damp_right", ctypes.c_float), ("damp_top", ctypes.c_float),
                ("damp_bottom", ctypes.c_float), ("damp_front", ctypes.c_float), ("damp_back", ctypes.c_float)]


class MonoSource(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]


class StereoListener(ctypes.Structure):
    """Head position, ear spacing (m), yaw / pitch / roll (degrees) and the
    head filters (StereoListener + HeadModel, TrueAudioVR.h:62-81)."""
    _fields_ = [("head_x", ctypes.c_float), ("head_y", ctypes.c_float), ("head_z", ctypes.c_float),
                ("ear_spacing", ctypes.c_float), ("yaw", ctypes.c_float), ("pitch", ctypes.c_float),
                ("roll", ctypes.c_float), ("filter_length", ctypes.c_int32),
                ("low_pass", ctypes.c_float * HEAD_FILTER_MAX), ("high_pass", ctypes.c_float * HEAD_FILTER_MAX)]


def make_room(width, height, length, damp=0.5, **walls):
    """RoomDefinition with every wall at `damp` unless named (damp_left=...)."""
    r = RoomDefinition(width, height, length, *([float(damp)] * 6))
    for k, v in walls.items():
        setattr(r, k, float(v))
    return r


def make_listener(head, ear_spacing=0.16, yaw=0.0, pitch=0.0, roll=0.0):
    """StereoListener with the simple head model of its ear spacing."""
    lst = StereoListener(*[float(v) for v in head], float(ear_spacing), float(yaw), float(pitch), float(roll))
    rc = generateSimpleHeadRelatedTransform(lst, ear_spacing)
    if rc != AMF_OK:
        raise ValueError(f"generateSimpleHeadRelatedTransform: AMF_RESULT {rc}")
    return lst


def generateSimpleHeadRelatedTransform(listener, ear_spacing):
    """Fills listener.filter_length / low_pass / high_pass
    (TrueAudioVR.cpp:361-415); returns an AMF_RESULT."""
    return _lib.load().tanx_room_head_model(float(ear_spacing), ctypes.byref(listener))


def _is_device_tensor(t):
    return hasattr(t, "is_cuda") and t.is_cuda


def GenerateRoomResponses(context, rooms, sources, listeners, sample_rate, response_length, out, flags=0,
                          max_bounces=0):
    """Responses k = 0..n-1 into rows 2k (left) and 2k+1 (right) of the
    float32 CUDA tensor `out` (rows >= 2n, row stride >= response_length).
    Asynchronous on the context's process stream."""
    n = len(rooms)
    if not (len(sources) == n and len(listeners) == n):
        raise ValueError("rooms, sources and listeners differ in count")
    if out.dim() != 2 or out.shape[0] < 2 * n or out.shape[1] < response_length or out.stride(1) != 1:
        raise ValueError("out must be a (>= 2n, >= response_length) row-major tensor")
    context.order_after_torch()
    ra = (RoomDefinition * n)(*rooms)
    sa = (MonoSource * n)(*sources)
    la = (StereoListener * n)(*listeners)
    return context._lib.tanx_room_response_device(context.handle, n, ra, sa, la, int(sample_rate),
                                                  int(response_length), int(flags), int(max_bounces),
                                                  out.data_ptr(), out.stride(0))


def generateRoomResponse(context, room, source, listener, sample_rate, response_length, responseL, responseR,
                         flags=0, max_bounces=0):
    """TrueAudioVRimpl::generateRoomResponse (TrueAudioVR.cpp:659-778):
    numpy float32 arrays are filled on return; CUDA tensors asynchronously
    (they are written through a (2, response_length) device buffer)."""
    lib = context._lib
    if _is_device_tensor(responseL):
        import torch
        tmp = torch.empty((2, response_length), dtype=torch.float32, device=responseL.device)
        rc = GenerateRoomResponses(context, [room], [source], [listener], sample_rate, response_length, tmp, flags,
                                   max_bounces)
        if rc == AMF_OK:
            proc = torch.cuda.ExternalStream(lib.tanx_ctx_process_stream(context.handle))
            with torch.cuda.stream(proc):
                responseL[:response_length].copy_(tmp[0])
                responseR[:response_length].copy_(tmp[1])
            tmp.record_stream(proc)  # the allocator keeps tmp until the copies ran
        return rc
    for a in (responseL, responseR):
        if not (isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous
                and a.size >= response_length):
            raise ValueError("host responses must be C-contiguous float32 arrays of response_length")
    return lib.tanx_room_response(context.handle, ctypes.byref(room), ctypes.byref(source), ctypes.byref(listener),
                                  int(sample_rate), int(response_length), int(flags), int(max_bounces),
                                  responseL.ctypes.data_as(_lib._f32p), responseR.ctypes.data_as(_lib._f32p))
