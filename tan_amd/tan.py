"""Python mirror of TAN's TANContext / TANConvolution / TANFFT interfaces.

Same class and method names, argument meaning and AMF_RESULT return codes as
tan/tanlibrary/include/TrueAudioNext.h:125-250 (TANConvolution), 493-541
(TANFFT) and 580-597 (TANContext).  Every call goes through the C ABI of the
HIP library (include/tan_mi355x.h); host arrays are numpy float32, device
arrays are torch tensors on the GPU (the stand-ins for the reference's
cl_mem overloads).
"""
import ctypes

import numpy as np

from . import _lib

# AMF_RESULT (amf/public/include/core/Result.h:43-122)
AMF_OK = 0
AMF_FAIL = 1
AMF_UNEXPECTED = 2
AMF_INVALID_ARG = 4
AMF_OUT_OF_MEMORY = 6
AMF_INVALID_POINTER = 7
AMF_NOT_IMPLEMENTED = 9
AMF_NOT_SUPPORTED = 10
AMF_NOT_FOUND = 11
AMF_ALREADY_INITIALIZED = 12
AMF_NOT_INITIALIZED = 13
AMF_WRONG_STATE = 15
AMF_NO_DEVICE = 17
AMF_INPUT_FULL = 25
AMF_TAN_UNSUPPORTED_VERSION = 43

# TAN_CONVOLUTION_METHOD (TrueAudioNext.h:81-98)
TAN_CONVOLUTION_METHOD_FFT_OVERLAP_ADD = 0
TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_UNIFORM = 1
TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM = 2
TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED = 3
TAN_CONVOLUTION_METHOD_FHT_UNIFORM_PARTITIONED = 4
TAN_CONVOLUTION_METHOD_FFT_UNIFORM_HEAD_TAIL = 5
TAN_CONVOLUTION_METHOD_FHT_UNIFORM_HEAD_TAIL = 6
TAN_CONVOLUTION_METHOD_TIME_DOMAIN = 7
TAN_CONVOLUTION_METHOD_FHT_NONUNIFORM_PARTITIONED = 8
TAN_CONVOLUTION_METHOD_FFT_NONUNIFORM_PARTITIONED = 9
TAN_CONVOLUTION_METHOD_USE_PROCESS_FINALIZE = 0x8000
TAN_CONVOLUTION_METHOD_USE_PROCESS_TAILTHREAD = 0xC000

TAN_CONVOLUTION_CHANNEL_FLAG_PROCESS = 0
TAN_CONVOLUTION_CHANNEL_FLAG_STOP_INPUT = 0x01
TAN_CONVOLUTION_CHANNEL_FLAG_FLUSH_STREAM = 0x02
TAN_CONVOLUTION_OPERATION_FLAG_NONE = 0
TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY = 0x01

# TAN_FFT_TRANSFORM_DIRECTION (TrueAudioNext.h:493-501)
TAN_FFT_TRANSFORM_DIRECTION_FORWARD = 0
TAN_FFT_TRANSFORM_DIRECTION_BACKWARD = 1
TAN_FFT_R2C_TRANSFORM_DIRECTION_FORWARD = 2
TAN_FFT_C2R_TRANSFORM_DIRECTION_BACKWARD = 3
TAN_FFT_R2C_PLANAR_TRANSFORM_DIRECTION_FORWARD = 4
TAN_FFT_C2R_PLANAR_TRANSFORM_DIRECTION_BACKWARD = 5

# IR spectrum storage precision (an option of this library)
TAN_IR_PRECISION_FP32 = 0
TAN_IR_PRECISION_FP16 = 1
TAN_IR_PRECISION_BF16 = 2

TAN_VERSION_MAJOR, TAN_VERSION_MINOR, TAN_VERSION_RELEASE, TAN_VERSION_BUILD = 1, 3, 1, 1
TAN_FULL_VERSION = ((TAN_VERSION_MAJOR << 48) | (TAN_VERSION_MINOR << 32)
                    | (TAN_VERSION_RELEASE << 16) | TAN_VERSION_BUILD)

_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def _is_device_tensor(x):
    try:
        import torch
    except ImportError:
        return False
    return isinstance(x, torch.Tensor) and x.is_cuda


def _host_rows(arrs, channels, length=None, writable=False):
    """Row-pointer array for host float32 data: a 2-D array or a list of 1-D arrays.

    Returns (pointer array, keepalive list)."""
    if isinstance(arrs, np.ndarray) and arrs.ndim == 2:
        rows = [arrs[i] for i in range(arrs.shape[0])]
    else:
        rows = list(arrs)
    if len(rows) < channels:
        raise ValueError("fewer channel arrays than channels")
    keep = []
    ptrs = (_f32p * channels)()
    for i in range(channels):
        r = rows[i]
        if not (isinstance(r, np.ndarray) and r.dtype == np.float32 and r.flags.c_contiguous):
            if writable:
                raise ValueError("output rows must be C-contiguous float32 numpy arrays")
            r = np.ascontiguousarray(r, dtype=np.float32)
        if length is not None and r.size < length:
            raise ValueError("channel array shorter than the requested length")
        keep.append(r)
        ptrs[i] = r.ctypes.data_as(_f32p)
    return ptrs, keep


def _flags_ptr(flags, channels):
    if flags is None:
        return None, None
    arr = np.ascontiguousarray(flags, dtype=np.uint32)
    if arr.size < channels:
        raise ValueError("flag array shorter than channel count")
    return arr.ctypes.data_as(_u32p), arr


class TANContext:
    """TANCreateContext + InitOpenCL equivalent: a HIP device with a process
    stream and an update stream (TANContextImpl.cpp:87-132, 264-328)."""

    def __init__(self, device=0, version=TAN_FULL_VERSION, conv_stream=None, general_stream=None):
        """conv_stream / general_stream: the caller's streams (torch.cuda.Stream
        or hipStream_t as int) for Process and IR updates, and for TANFFT, as
        the reference's InitOpenCL(convolutionQueue, generalQueue) takes the
        application's queues; None: a stream of the context's own.  The
        context keeps references to the stream objects it is given."""
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        self._ext = (conv_stream, general_stream)
        if conv_stream is None and general_stream is None:
            rc = self._lib.tanx_ctx_create(version, device, ctypes.byref(h))
        else:
            def raw(st):
                return None if st is None else int(getattr(st, "cuda_stream", st))
            rc = self._lib.tanx_ctx_create_on_streams(version, device, raw(conv_stream), raw(general_stream),
                                                      ctypes.byref(h))
        if rc != AMF_OK:
            raise RuntimeError(f"TANCreateContext failed: AMF_RESULT {rc}")
        self.handle = h
        self.device = device
        self._streams = None
        self._raw = (self._lib.tanx_ctx_process_stream(h) or 0, self._lib.tanx_ctx_update_stream(h) or 0)

    def order_after_torch(self, update=False):
        """Makes the work this context queues next (on its process stream,
        and with update=True on its general stream too) start after the work
        queued so far on torch's current stream, which may be writing the
        tensors the next call reads or writes.  The library's calls are
        asynchronous on its own streams; the device-tensor methods of this
        module call this first."""
        import torch
        if self._streams is None:
            self._streams = (torch.cuda.ExternalStream(self._raw[0]), torch.cuda.ExternalStream(self._raw[1]))
        cur = torch.cuda.current_stream()
        # work queued on the same stream is ordered already
        if cur.cuda_stream != self._raw[0]:
            self._streams[0].wait_stream(cur)
        if update and cur.cuda_stream != self._raw[1]:
            self._streams[1].wait_stream(cur)

    def SetGroupQueue(self, enable=True):
        """Group calls of this context as work-queue launches (the device's
        resident workgroups take the launch's items from a counter)."""
        return self._lib.tanx_ctx_set_group_queue(self.handle, int(bool(enable)))

    def EnableGroupTiming(self, enable=True):
        """Launch-stamped timing of this context's group launches."""
        return self._lib.tanx_ctx_enable_group_timing(self.handle, int(bool(enable)))

    def ReadGroupTiming(self, reset=False):
        """{tick_ms, tick_launches, tick_alg_bytes, frame_ms, frame_launches,
        frame_alg_bytes, objects} of the group launches since the last reset."""
        t = _lib.KernelTiming()
        rc = self._lib.tanx_ctx_read_group_timing(self.handle, ctypes.byref(t), int(bool(reset)))
        if rc != AMF_OK:
            raise RuntimeError(f"read_group_timing failed: {rc}")
        return {"tick_ms": t.head_ms, "tick_launches": t.head_launches, "tick_alg_bytes": t.head_alg_bytes,
                "frame_ms": t.tail_ms, "frame_launches": t.tail_launches, "frame_alg_bytes": t.tail_alg_bytes,
                "objects": t.blocks}

    def Terminate(self):
        if self.handle:
            self._streams = None
            rc = self._lib.tanx_ctx_destroy(self.handle)
            self.handle = None
            return rc
        return AMF_OK

    def __del__(self):
        try:
            self.Terminate()
        except Exception:
            pass


class TANConvolution:
    """TANConvolution on the GPU engine (ConvolutionImpl.cpp)."""

    def __init__(self, context, ir_precision=TAN_IR_PRECISION_FP32, nu_multiple=0, nu_phase=0, pass_blocks=0,
                 pass_phase=0, max_tail_chunks=0, ir_slots=0):
        """ir_precision: storage of the IR spectra (TAN_IR_PRECISION_*);
        nu_multiple: blocks per frame for the non-uniform methods (0: chosen
        by the library, 1: uniform partitions); nu_phase: frame phase in
        blocks, modulo the multiple (the first frame ends at block
        multiple - 1 - phase), so that objects advanced together can spread
        their frame work over the ticks; pass_blocks: blocks whose partial
        sums one tail pass computes for uniform objects (1..16, 0: chosen by
        the library, 1: a tail every block); pass_phase: uniform objects with
        multi-block passes run their passes in the launches of blocks b with
        (b + pass_phase) % pass_blocks == 0 (after block 0), so objects
        advanced together with phases 0..pass_blocks-1 spread their passes
        over the ticks; max_tail_chunks: cap on the partition chunks per
        channel of a pass (0: the library's choice), which sets the partial
        rows kept per block; ir_slots: 2 keeps two IR spectrum slots instead
        of 3 (uniform objects; an update issued while a switch is armed then
        returns AMF_INPUT_FULL)."""
        self._ctx = context
        self._lib = context._lib
        self.handle = None
        self.channels = 0
        self.block = 0
        self.ir_precision = ir_precision
        self.nu_multiple = nu_multiple
        self.nu_phase = nu_phase
        self.pass_blocks = pass_blocks
        self.pass_phase = pass_phase
        self.max_tail_chunks = max_tail_chunks
        self.ir_slots = ir_slots

    # -- Init family (ConvolutionImpl.cpp:179-305, 1113-1259) --
    def Init(self, method, response_length, buffer_size, channels):
        return self.InitGpu(method, response_length, buffer_size, channels)

    def InitGpu(self, method, response_length, buffer_size, channels):
        if self.handle:
            return AMF_ALREADY_INITIALIZED
        h = ctypes.c_void_p()
        opt = _lib.ConvOptions()
        opt.ir_precision = self.ir_precision
        opt.nu_multiple = self.nu_multiple
        opt.nu_phase = self.nu_phase
        opt.pass_blocks = self.pass_blocks
        opt.pass_phase = self.pass_phase
        opt.max_tail_chunks = self.max_tail_chunks
        opt.ir_slots = self.ir_slots
        rc = self._lib.tanx_conv_create_ex(self._ctx.handle, method, response_length, buffer_size,
                                           channels, ctypes.addressof(opt), ctypes.byref(h))
        if rc == AMF_OK:
            self.handle = h
            self.channels = channels
            self.block = buffer_size
        return rc

    def InitCpu(self, method, response_length, buffer_size, channels):
        # the CPU engine of the reference is not part of this library
        return AMF_NOT_SUPPORTED

    def Terminate(self):
        if self.handle:
            rc = self._lib.tanx_conv_destroy(self.handle)
            self.handle = None
            return rc
        return AMF_OK

    def GetContext(self):
        return self._ctx

    def __del__(self):
        try:
            self.Terminate()
        except Exception:
            pass

    # -- responses --
    def UpdateResponseTD(self, buffers, num_samples=None, flag_masks=None, operation_flags=0):
        if not self.handle:
            return AMF_NOT_INITIALIZED
        fp, _fk = _flags_ptr(flag_masks, self.channels)
        if _is_device_tensor(buffers):
            t = buffers
            if num_samples is None:
                num_samples = t.shape[1]
            # the library reads num_samples per row through the pointer and
            # row stride: a shorter view would be read past its end
            if t.dim() != 2 or t.shape[0] < self.channels or t.shape[1] < num_samples:
                return AMF_INVALID_ARG
            # the transform runs on the process stream
            self._ctx.order_after_torch()
            return self._lib.tanx_conv_update_td_device(self.handle, t.data_ptr(), t.stride(0),
                                                        num_samples, fp, operation_flags)
        if num_samples is None:
            num_samples = (buffers.shape[1] if isinstance(buffers, np.ndarray)
                           else min(len(b) for b in buffers))
        ptrs, _keep = _host_rows(buffers, self.channels, num_samples)
        return self._lib.tanx_conv_update_td(self.handle, ptrs, num_samples, fp, operation_flags)

    def UpdateResponseFD(self, *args, **kwargs):
        # not implemented in the reference either (ConvolutionImpl.cpp:697-717)
        return AMF_NOT_SUPPORTED if self.handle else AMF_NOT_INITIALIZED

    # -- processing --
    def Process(self, inputs, outputs, num_samples=None, flag_masks=None):
        """Returns (AMF_RESULT, samples processed)."""
        if not self.handle:
            return AMF_NOT_INITIALIZED, 0
        fp, _fk = _flags_ptr(flag_masks, self.channels)
        done = ctypes.c_size_t(0)
        if _is_device_tensor(inputs):
            if num_samples is None:
                num_samples = inputs.shape[1]
            # whole blocks are read from and written to every row through the
            # pointer and row stride: views shorter than that would be overrun
            nwhole = (num_samples // self.block) * self.block
            for t in (inputs, outputs):
                if not _is_device_tensor(t) or t.dim() != 2 or t.shape[0] < self.channels or t.shape[1] < nwhole:
                    return AMF_INVALID_ARG, 0
            self._ctx.order_after_torch()
            rc = self._lib.tanx_conv_process_device(
                self.handle, inputs.data_ptr(), inputs.stride(0), outputs.data_ptr(),
                outputs.stride(0), num_samples, fp, ctypes.byref(done))
            return rc, done.value
        if num_samples is None:
            num_samples = (inputs.shape[1] if isinstance(inputs, np.ndarray)
                           else min(len(b) for b in inputs))
        pin, _ki = _host_rows(inputs, self.channels, min(num_samples, self.block))
        pout, _ko = _host_rows(outputs, self.channels, min(num_samples, self.block), writable=True)
        rc = self._lib.tanx_conv_process(self.handle, pin, pout, num_samples, fp, ctypes.byref(done))
        return rc, done.value

    def ProcessDirect(self, impulse_responses, inputs, outputs, num_samples=None, nz_first_last=None):
        """ProcessDirect (ConvolutionImpl.cpp:719-762): TIME_DOMAIN objects
        only.  Host arrays (channels, >= n), or CUDA tensors (channels, >= n
        samples; IRs (channels, ir_len)), which run on the context's process
        stream; nz_first_last: None or (channels, 2) first / last+1 non-zero
        taps.  Returns (AMF_RESULT, processed)."""
        if not self.handle:
            return AMF_NOT_INITIALIZED, 0
        if _is_device_tensor(inputs):
            if num_samples is None:
                num_samples = inputs.shape[1]
            n = min(num_samples, self.Info()["ir_len"])
            views = ((impulse_responses, self.Info()["ir_len"]), (inputs, n), (outputs, n))
            for t, need in views:
                if (not _is_device_tensor(t) or t.dim() != 2 or t.shape[0] < self.channels or t.shape[1] < need
                        or t.dtype.itemsize != 4 or t.stride(1) != 1):
                    return AMF_INVALID_ARG, 0
            nz = None
            if nz_first_last is not None:
                nz = np.ascontiguousarray(nz_first_last, dtype=np.int32).reshape(-1)
            self._ctx.order_after_torch()
            done = ctypes.c_size_t(0)
            rc = self._lib.tanx_conv_process_direct_device(
                self.handle, impulse_responses.data_ptr(), impulse_responses.stride(0), inputs.data_ptr(),
                inputs.stride(0), outputs.data_ptr(), outputs.stride(0), num_samples, ctypes.byref(done),
                nz.ctypes.data if nz is not None else None)
            return rc, done.value
        if num_samples is None:
            num_samples = inputs.shape[1] if isinstance(inputs, np.ndarray) else min(len(b) for b in inputs)
        pir, _k1 = _host_rows(impulse_responses, self.channels)
        pin, _k2 = _host_rows(inputs, self.channels, num_samples)
        pout, _k3 = _host_rows(outputs, self.channels, num_samples, writable=True)
        nz = None
        if nz_first_last is not None:
            nz = np.ascontiguousarray(nz_first_last, dtype=np.int32).reshape(-1)
        done = ctypes.c_size_t(0)
        rc = self._lib.tanx_conv_process_direct(self.handle, pir, pin, pout, num_samples, ctypes.byref(done),
                                                nz.ctypes.data if nz is not None else None)
        return rc, done.value

    def ProcessFinalize(self):
        if not self.handle:
            return AMF_NOT_INITIALIZED
        return self._lib.tanx_conv_finalize(self.handle)

    def GetNextFreeChannel(self, flag_masks):
        """Returns (AMF_RESULT, channel index)."""
        if not self.handle:
            return AMF_NOT_INITIALIZED, None
        fp, _fk = _flags_ptr(flag_masks, self.channels)
        idx = ctypes.c_uint32(0)
        rc = self._lib.tanx_conv_get_next_free_channel(self.handle, ctypes.byref(idx), fp)
        return rc, (idx.value if rc == AMF_OK else None)

    # -- extras of this implementation --
    def Flush(self, channel):
        return self._lib.tanx_conv_flush(self.handle, channel)

    def Synchronize(self):
        return self._lib.tanx_conv_synchronize(self.handle)

    def Info(self):
        info = _lib.ConvInfo()
        self._lib.tanx_conv_get_info(self.handle, ctypes.byref(info))
        return {k: getattr(info, k) for k, _ in info._fields_}

    def SetTailChunks(self, chunks):
        return self._lib.tanx_conv_set_tail_chunks(self.handle, chunks)

    def Autotune(self):
        """Returns (AMF_RESULT, chosen tail chunks, microseconds per block)."""
        n = ctypes.c_uint32(0)
        us = ctypes.c_double(0.0)
        rc = self._lib.tanx_conv_autotune(self.handle, ctypes.byref(n), ctypes.byref(us))
        return rc, n.value, us.value

    def SetGraph(self, use_graph=True):
        return self._lib.tanx_conv_set_graph(self.handle, int(bool(use_graph)))

    def SetFused(self, fused=True):
        return self._lib.tanx_conv_set_fused(self.handle, int(bool(fused)))

    def EnableTiming(self, enable=True):
        return self._lib.tanx_conv_enable_timing(self.handle, int(bool(enable)))

    def ReadTiming(self, reset=False):
        t = _lib.KernelTiming()
        rc = self._lib.tanx_conv_read_timing(self.handle, ctypes.byref(t), int(bool(reset)))
        if rc != AMF_OK:
            raise RuntimeError(f"read_timing failed: {rc}")
        return {k: getattr(t, k) for k, _ in t._fields_}

    def LastWasCrossfade(self):
        return bool(self._lib.tanx_conv_last_was_crossfade(self.handle))

    def ProcessStream(self):
        return self._lib.tanx_conv_process_stream(self.handle)


class TANFFT:
    """TANFFT on the GPU (FFTImpl.cpp:507-575)."""

    def __init__(self, context):
        self._ctx = context
        self._lib = context._lib
        self._init = False

    def Init(self):
        self._init = True
        return AMF_OK

    def Terminate(self):
        self._init = False
        return AMF_OK

    def GetContext(self):
        return self._ctx

    def Transform(self, direction, log2len, channels, inputs, outputs):
        if not self._init:
            return AMF_NOT_INITIALIZED
        if _is_device_tensor(inputs):
            # one channel stride for both; every row holds the 2^(log2len+1) floats
            n = 2 << log2len if 0 < log2len < 31 else 0
            for t in (inputs, outputs):
                if (not _is_device_tensor(t) or t.dim() != 2 or t.dtype.itemsize != 4 or t.stride(1) != 1
                        or t.shape[0] < channels or t.shape[1] < n or t.stride(0) != inputs.stride(0)):
                    return AMF_INVALID_ARG
            self._ctx.order_after_torch(update=True)
            return self._lib.tanx_fft_transform_device(self._ctx.handle, direction, log2len, channels,
                                                       inputs.data_ptr(), outputs.data_ptr(),
                                                       inputs.stride(0))
        n = 2 << log2len
        pin, _ki = _host_rows(inputs, channels, n)
        pout, _ko = _host_rows(outputs, channels, n, writable=True)
        return self._lib.tanx_fft_transform(self._ctx.handle, direction, log2len, channels, pin, pout)

    def TransformR2CMatrixCores(self, log2len, channels, inputs, outputs, split=3):
        """Real-to-complex DFT of 2^log2len samples per channel as a bf16 GEMM on
        the matrix cores (split = 3: compensated high/low products).  torch
        CUDA tensors (channels, >= n) in and (channels, >= n + 2) out, bins as
        (re, im) pairs; asynchronous on the context's process stream."""
        self._ctx.order_after_torch()
        return self._lib.tanx_dft_r2c_mfma_device(self._ctx.handle, log2len, channels, inputs.data_ptr(),
                                                  inputs.stride(0), outputs.data_ptr(), outputs.stride(0), split)


def TANCreateContext(version=TAN_FULL_VERSION, device=0):
    """Returns (AMF_RESULT, TANContext or None), as TANCreateContext does."""
    try:
        return AMF_OK, TANContext(device=device, version=version)
    except RuntimeError as e:
        msg = str(e)
        code = int(msg.rsplit(" ", 1)[-1]) if msg[-1].isdigit() else AMF_FAIL
        return code, None


def TANCreateConvolution(context):
    return AMF_OK, TANConvolution(context)


def TANCreateFFT(context):
    return AMF_OK, TANFFT(context)


def GroupProcess(convs, inputs, outputs, flag_masks=None):
    """One block for each convolution object in one kernel launch.

    convs: TANConvolution objects of one context; inputs/outputs: torch CUDA
    tensors (channels, >= block) per object; flag_masks: None or a list of
    per-object flag arrays (None entries allowed).  Asynchronous on the
    context's process stream, like device-memory Process."""
    n = len(convs)
    if n == 0:
        return AMF_OK
    for c, i, o in zip(convs, inputs, outputs):
        for t in (i, o):
            if not _is_device_tensor(t) or t.dim() != 2 or t.shape[0] < c.channels or t.shape[1] < c.block:
                return AMF_INVALID_ARG
    lib = convs[0]._lib
    convs[0]._ctx.order_after_torch()
    hs = (ctypes.c_void_p * n)(*[c.handle for c in convs])
    ins = (ctypes.c_void_p * n)(*[t.data_ptr() for t in inputs])
    outs = (ctypes.c_void_p * n)(*[t.data_ptr() for t in outputs])
    ist = (ctypes.c_size_t * n)(*[t.stride(0) for t in inputs])
    ost = (ctypes.c_size_t * n)(*[t.stride(0) for t in outputs])
    fptr = None
    keep = []
    if flag_masks is not None:
        arr = (ctypes.c_void_p * n)()
        for i, f in enumerate(flag_masks):
            if f is not None:
                a = np.ascontiguousarray(f, dtype=np.uint32)
                keep.append(a)
                arr[i] = a.ctypes.data
        fptr = ctypes.cast(arr, ctypes.c_void_p)
    return lib.tanx_group_process_device(n, hs, ins, ist, outs, ost, fptr)


def GroupUpdateResponseTD(convs, irs, lengths=None, flag_masks=None, operation_flags=0):
    """UpdateResponseTD of several objects of one context in one call
    (tanx_group_update_td_device): object i's update is the one
    convs[i].UpdateResponseTD(irs[i], lengths[i]) would make, but the IR
    transforms of all of them run in one launch per partition size.  irs:
    torch CUDA tensors (channels, >= length); lengths: None (each tensor's
    row length, as UpdateResponseTD takes it) or per-object sample counts."""
    n = len(convs)
    if n == 0:
        return AMF_OK
    lens = [h.shape[1] if lengths is None else int(lengths[i]) for i, h in enumerate(irs)]
    for c, h, ln in zip(convs, irs, lens):
        if (not _is_device_tensor(h) or h.dim() != 2 or h.shape[0] < c.channels or h.shape[1] < ln
                or h.dtype.itemsize != 4 or h.stride(1) != 1):
            return AMF_INVALID_ARG
    lib = convs[0]._lib
    convs[0]._ctx.order_after_torch()
    hs = (ctypes.c_void_p * n)(*[c.handle for c in convs])
    ptrs = (ctypes.c_void_p * n)(*[h.data_ptr() for h in irs])
    strides = (ctypes.c_size_t * n)(*[h.stride(0) for h in irs])
    ln = (ctypes.c_size_t * n)(*lens)
    fptr = None
    keep = []
    if flag_masks is not None:
        arr = (ctypes.c_void_p * n)()
        for i, f in enumerate(flag_masks):
            if f is not None:
                a = np.ascontiguousarray(f, dtype=np.uint32)
                keep.append(a)
                arr[i] = a.ctypes.data
        fptr = ctypes.cast(arr, ctypes.c_void_p)
    return lib.tanx_group_update_td_device(n, hs, ptrs, strides, ln, fptr, operation_flags)


# TAN_SAMPLE_TYPE (TrueAudioNext.h:73-77)
TAN_SAMPLE_TYPE_FLOAT = 0
TAN_SAMPLE_TYPE_SHORT = 1


def _sample_type(dtype):
    if dtype in (np.float32, "float32") or str(dtype).endswith("float32"):
        return TAN_SAMPLE_TYPE_FLOAT
    if dtype in (np.int16, "int16") or str(dtype).endswith("int16"):
        return TAN_SAMPLE_TYPE_SHORT
    raise ValueError(f"unsupported sample type {dtype}")


class TANConverter:
    """TANConverter (converter/ConverterImpl.cpp) on the GPU: float <-> short
    with gain and clamping, strided and offset."""

    def __init__(self, context):
        self._ctx = context
        self._lib = context._lib

    def Init(self):
        return AMF_OK

    def Terminate(self):
        return AMF_OK

    def GetContext(self):
        return self._ctx

    def Convert(self, inputs, input_step, num_samples, outputs, output_step, gain,
                input_offset=0, output_offset=0, want_clipped=True):
        """Returns (AMF_RESULT, clipped).  numpy arrays (float32 / int16) use the
        host overload; torch CUDA tensors the device overload (asynchronous
        unless want_clipped)."""
        clipped = ctypes.c_int32(0)
        cp = ctypes.byref(clipped) if want_clipped else None
        if _is_device_tensor(inputs):
            self._ctx.order_after_torch()
            rc = self._lib.tanx_convert_device(
                self._ctx.handle, inputs.data_ptr(), input_step, input_offset, _sample_type(inputs.dtype),
                outputs.data_ptr(), output_step, output_offset, _sample_type(outputs.dtype), num_samples,
                float(gain), cp)
            return rc, bool(clipped.value)
        if not (inputs.flags.c_contiguous and outputs.flags.c_contiguous):
            raise ValueError("host buffers must be C-contiguous")
        rc = self._lib.tanx_convert(
            self._ctx.handle, inputs.ctypes.data + input_offset * inputs.itemsize, input_step,
            _sample_type(inputs.dtype), outputs.ctypes.data + output_offset * outputs.itemsize, output_step,
            _sample_type(outputs.dtype), num_samples, float(gain), cp)
        return rc, bool(clipped.value)


class TANMixer:
    """TANMixer (mixer/MixerImpl.cpp) on the GPU: the channel sum in channel
    order."""

    def __init__(self, context):
        self._ctx = context
        self._lib = context._lib
        self.handle = None
        self.buffer_size = 0
        self.channels = 0

    def Init(self, buffer_size, num_channels):
        if self.handle:
            return AMF_ALREADY_INITIALIZED
        h = ctypes.c_void_p()
        rc = self._lib.tanx_mixer_create(self._ctx.handle, buffer_size, num_channels, ctypes.byref(h))
        if rc == AMF_OK:
            self.handle = h
            self.buffer_size = buffer_size
            self.channels = num_channels
        return rc

    def Terminate(self):
        if self.handle:
            rc = self._lib.tanx_mixer_destroy(self.handle)
            self.handle = None
            return rc
        return AMF_OK

    def GetContext(self):
        return self._ctx

    def __del__(self):
        try:
            self.Terminate()
        except Exception:
            pass

    def Mix(self, inputs, output, input_stride=None):
        """inputs: a (channels, >= buffer_size) array or a list of per-channel
        arrays; numpy -> host overload, torch CUDA -> device overloads (one
        2-D tensor: the strided one; a list: the pointer-list one)."""
        if not self.handle:
            return AMF_NOT_INITIALIZED
        if _is_device_tensor(output):
            self._ctx.order_after_torch()
            if _is_device_tensor(inputs):
                stride = inputs.stride(0) if input_stride is None else input_stride
                return self._lib.tanx_mix_device_strided(self.handle, inputs.data_ptr(), stride, output.data_ptr())
            arr = (ctypes.c_void_p * self.channels)(*[t.data_ptr() for t in inputs])
            return self._lib.tanx_mix_device(self.handle, arr, output.data_ptr())
        ptrs, _keep = _host_rows(inputs, self.channels, self.buffer_size)
        if not (isinstance(output, np.ndarray) and output.dtype == np.float32 and output.flags.c_contiguous):
            raise ValueError("output must be a C-contiguous float32 numpy array")
        return self._lib.tanx_mix(self.handle, ptrs, output.ctypes.data)


def TANCreateConverter(context):
    return AMF_OK, TANConverter(context)


def TANCreateMixer(context):
    return AMF_OK, TANMixer(context)
