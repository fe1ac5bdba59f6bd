"""Reduced-precision IR spectra (FP16 / BF16 with a per-partition max-abs
scale, FP32 accumulation).  The SNR these storage types allow is set by
their mantissas (about 71 dB for FP16's 11 bits and 53 dB for BF16's 8 bits
on this IR model); the FP32 path stays the drop-in at >= 120 dB."""
import numpy as np
import pytest

import tan_amd as T
from helpers import crossfade_model, snr_db

pytestmark = pytest.mark.gpu
UNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
BLOCKING = T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY
# thresholds: 6 dB under the mantissa-limited SNR
MIN_SNR = {T.TAN_IR_PRECISION_FP32: 120.0, T.TAN_IR_PRECISION_FP16: 65.0, T.TAN_IR_PRECISION_BF16: 47.0}


def _run(ctx, prec, L, B, C, nb, ups, x):
    conv = T.TANConvolution(ctx, ir_precision=prec)
    assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
    assert conv.Info()["ir_precision"] == prec
    y = np.zeros_like(x)
    xf = []
    for b in range(nb):
        if b in ups:
            assert conv.UpdateResponseTD(ups[b], L, None, BLOCKING) == T.AMF_OK
        out = np.zeros((C, B), np.float32)
        rc, done = conv.Process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]), out)
        assert rc == T.AMF_OK and done == B
        y[:, b * B:(b + 1) * B] = out
        if conv.LastWasCrossfade():
            xf.append(b)
    conv.Terminate()
    return y, xf


def test_precision_c4_shape(gpu_ctx):
    """B=256, L=131,072 (P=512), 4 channels, 560 blocks (every partition in
    use), one IR switch; SNR against the FP64 model and against the FP32 path."""
    L, B, C, nb = 131072, 256, 4, 560
    ups = {0: T.signals.ir_bank(C, L, 0), 300: T.signals.ir_bank(C, L, 1)}
    x = T.signals.input_bank(C, nb * B)
    outs = {}
    for prec in (T.TAN_IR_PRECISION_FP32, T.TAN_IR_PRECISION_FP16, T.TAN_IR_PRECISION_BF16):
        y, xf = _run(gpu_ctx, prec, L, B, C, nb, ups, x)
        assert xf == [2, 302]
        outs[prec] = y
    for c in range(C):
        ref = crossfade_model(x[c], [ups[0][c], ups[300][c]], [2, 302], B)
        for prec, y in outs.items():
            snr = snr_db(ref, y[c])
            snr32 = snr_db(outs[T.TAN_IR_PRECISION_FP32][c], y[c])
            print(f"ch{c} precision {prec}: SNR vs FP64 {snr:.1f} dB, vs FP32 path {snr32:.1f} dB", flush=True)
            assert snr >= MIN_SNR[prec]


def test_precision_short_block_sizes(gpu_ctx):
    for B in (128, 512):
        L, C, nb = 8192, 2, 40
        ups = {0: T.signals.ir_bank(C, L, 3)}
        x = T.signals.input_bank(C, nb * B)
        y, _ = _run(gpu_ctx, T.TAN_IR_PRECISION_FP16, L, B, C, nb, ups, x)
        for c in range(C):
            ref = crossfade_model(x[c], [ups[0][c]], [2], B)
            assert snr_db(ref, y[c]) >= MIN_SNR[T.TAN_IR_PRECISION_FP16]


def test_precision_partial_update_keeps_rows_and_scales(gpu_ctx):
    """A channel skipped by an update keeps its FP16 rows and scales."""
    L, B, C, nb = 4096, 128, 3, 40
    h0, h1 = T.signals.ir_bank(C, L, 0), T.signals.ir_bank(C, L, 1)
    x = T.signals.input_bank(C, nb * B)
    conv = T.TANConvolution(gpu_ctx, ir_precision=T.TAN_IR_PRECISION_FP16)
    conv.InitGpu(UNIFORM, L, B, C)
    skip = np.array([0, 1, 0], np.uint32)
    y = np.zeros_like(x)
    for b in range(nb):
        if b == 0:
            conv.UpdateResponseTD(h0, L, None, BLOCKING)
        if b == 15:
            conv.UpdateResponseTD(h1, L, skip, BLOCKING)
        out = np.zeros((C, B), np.float32)
        conv.Process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]), out)
        y[:, b * B:(b + 1) * B] = out
    conv.Terminate()
    ref1 = crossfade_model(x[1], [h0[1], h0[1]], [2, 17], B)
    ref0 = crossfade_model(x[0], [h0[0], h1[0]], [2, 17], B)
    assert snr_db(ref1, y[1]) >= 65.0 and snr_db(ref0, y[0]) >= 65.0
