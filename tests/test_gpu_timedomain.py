"""TAN_CONVOLUTION_METHOD_TIME_DOMAIN: direct convolution for short (HRTF-
length) responses with the same switch schedule as the uniform engine, and
ProcessDirect with per-call responses and non-zero tap ranges
(ConvolutionImpl.cpp:719-762, 2647-2782).  Pinned to the FP64 convolution."""
import numpy as np
import pytest

import tan_amd as T
from helpers import crossfade_model, fft_conv_f64, snr_db

pytestmark = pytest.mark.gpu
TD = T.TAN_CONVOLUTION_METHOD_TIME_DOMAIN
UNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
BLOCKING = T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY


def _run(ctx, method, L, B, C, x, irs, ups, device=True):
    import torch
    conv = T.TANConvolution(ctx)
    assert conv.InitGpu(method, L, B, C) == T.AMF_OK
    nb = x.shape[1] // B
    y = np.zeros_like(x)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    k = 0
    for b in range(nb):
        while k < len(ups) and ups[k] == b:
            assert conv.UpdateResponseTD(irs[k], L, None, BLOCKING if b == 0 else 0) == T.AMF_OK
            k += 1
        if device:
            rc, done = conv.Process(xd[:, b * B:], yd[:, b * B:], B)
        else:
            rc, done = conv.Process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]), y[:, b * B:(b + 1) * B], B)
        assert rc == T.AMF_OK and done == B
    conv.Synchronize()
    conv.Terminate()
    return yd.cpu().numpy() if device else y


@pytest.mark.parametrize("L,B,C", [(128, 64, 3), (500, 256, 2), (1024, 1024, 1), (4096, 256, 2), (200, 128, 4)])
def test_time_domain_matches_fp64_with_updates(gpu_ctx, L, B, C):
    nb = 40
    x = T.signals.input_bank(C, nb * B, base_seed=31)
    ups = [0, 9, 22, 24]  # 24: issued while the switch of 22 (block 24) is armed
    irs = [T.signals.ir_bank(C, L, k, base_seed=32) for k in range(len(ups))]
    y = _run(gpu_ctx, TD, L, B, C, x, irs, ups)
    # two blocks after each update; the one issued while armed is handed off
    # by block 25, armed at 26 and crossfades at 27
    sw = [2, 11, 24, 27]
    for c in range(C):
        ref = crossfade_model(x[c], [h[c] for h in irs], sw, B)
        assert np.abs(y[c] - ref).max() <= 1e-5, f"channel {c}"
        assert snr_db(ref, y[c]) >= 120.0


def test_time_domain_equals_uniform_engine(gpu_ctx):
    L, B, C = 512, 128, 3
    x = T.signals.input_bank(C, 30 * B, base_seed=41)
    irs = [T.signals.ir_bank(C, L, k, base_seed=42) for k in range(2)]
    yt = _run(gpu_ctx, TD, L, B, C, x, irs, [0, 12])
    yu = _run(gpu_ctx, UNIFORM, L, B, C, x, irs, [0, 12])
    assert np.abs(yt - yu).max() <= 1e-5
    yh = _run(gpu_ctx, TD, L, B, C, x, irs, [0, 12], device=False)
    np.testing.assert_array_equal(yt, yh)


def test_process_direct_with_tap_ranges(gpu_ctx):
    L, B, C = 1024, 256, 2
    conv = T.TANConvolution(gpu_ctx)
    assert conv.InitGpu(TD, L, B, C) == T.AMF_OK
    x = T.signals.input_bank(C, 3000, base_seed=51)
    ir = T.signals.ir_bank(C, L, 0, base_seed=52)
    nz = np.array([[10, 700], [0, 1024]], np.int32)
    y = np.zeros_like(x)
    done_total = 0
    for off, n in ((0, 1000), (1000, 1024), (2024, 976)):
        out = np.zeros((C, n), np.float32)
        rc, done = conv.ProcessDirect(ir, np.ascontiguousarray(x[:, off:off + n]), out, n, nz)
        assert rc == T.AMF_OK and done == n
        y[:, off:off + n] = out
        done_total += done
    for c in range(C):
        h = np.zeros(L, np.float32)
        h[nz[c, 0]:nz[c, 1]] = ir[c, nz[c, 0]:nz[c, 1]]
        ref = fft_conv_f64(x[c], h)
        assert np.abs(y[c] - ref).max() <= 1e-5
    # at most the IR length per call (ConvolutionImpl.cpp:733-734)
    big = np.zeros((C, 2000), np.float32)
    rc, done = conv.ProcessDirect(ir, big, big.copy(), 2000)
    assert rc == T.AMF_OK and done == L
    conv.Terminate()
    u = T.TANConvolution(gpu_ctx)
    assert u.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
    assert u.ProcessDirect(ir, x[:, :B], np.zeros((C, B), np.float32), B)[0] == T.AMF_NOT_SUPPORTED
    u.Terminate()


def test_process_direct_device_equals_host(gpu_ctx):
    """ProcessDirect on device tensors (tanx_conv_process_direct_device)
    gives the host overload's bits, call after call (the history carries
    over), with tap ranges; the caller's IR rows are read in place."""
    import torch
    L, B, C = 1024, 256, 3
    x = T.signals.input_bank(C, 3000, base_seed=53)
    ir = T.signals.ir_bank(C, L, 0, base_seed=54)
    nz = np.array([[10, 700], [0, 1024], [300, 301]], np.int32)
    host = T.TANConvolution(gpu_ctx)
    dev = T.TANConvolution(gpu_ctx)
    assert host.InitGpu(TD, L, B, C) == T.AMF_OK and dev.InitGpu(TD, L, B, C) == T.AMF_OK
    # IR rows with a stride wider than the IR, inputs and outputs as views
    ird = torch.zeros((C, L + 64), dtype=torch.float32, device="cuda")
    ird[:, :L] = torch.from_numpy(ir).cuda()
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    yh = np.zeros_like(x)
    for off, n in ((0, 1000), (1000, 1024), (2024, 976)):
        out = np.zeros((C, n), np.float32)
        rc, done = host.ProcessDirect(ir, np.ascontiguousarray(x[:, off:off + n]), out, n, nz)
        assert rc == T.AMF_OK and done == n
        yh[:, off:off + n] = out
        rc, done = dev.ProcessDirect(ird, xd[:, off:], yd[:, off:], n, nz)
        assert rc == T.AMF_OK and done == n
    assert dev.Synchronize() == T.AMF_OK
    np.testing.assert_array_equal(yd.cpu().numpy(), yh)
    # views shorter than the IR or the samples are refused before any launch
    assert dev.ProcessDirect(ird[:, :L - 1], xd, yd, 100)[0] == T.AMF_INVALID_ARG
    assert dev.ProcessDirect(ird, xd[:, :50], yd, 100)[0] == T.AMF_INVALID_ARG
    host.Terminate()
    dev.Terminate()


def test_time_domain_flush_and_group_rejection(gpu_ctx):
    import torch
    L, B, C = 256, 64, 2
    conv = T.TANConvolution(gpu_ctx)
    assert conv.InitGpu(TD, L, B, C) == T.AMF_OK
    ir = T.signals.ir_bank(C, L, 0, base_seed=61)
    assert conv.UpdateResponseTD(ir, L, None, BLOCKING) == T.AMF_OK
    x = T.signals.input_bank(C, 20 * B, base_seed=62)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    for b in range(20):
        flags = np.array([T.TAN_CONVOLUTION_CHANNEL_FLAG_FLUSH_STREAM, 0], np.uint32) if b == 10 else None
        rc, _ = conv.Process(xd[:, b * B:], yd[:, b * B:], B, flags)
        assert rc == T.AMF_OK
    conv.Synchronize()
    y = yd.cpu().numpy()
    assert np.abs(y[0, 10 * B:] - fft_conv_f64(x[0, 10 * B:], ir[0])).max() <= 1e-5
    assert T.GroupProcess([conv], [xd[:, :B]], [yd[:, :B]]) == T.AMF_NOT_SUPPORTED
    assert T.TANConvolution(gpu_ctx).InitGpu(TD, 8192, B, 1) == T.AMF_INVALID_ARG  # beyond 4096 taps
    conv.Terminate()
