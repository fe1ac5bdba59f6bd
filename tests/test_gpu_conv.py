"""GPU parity of the convolution path against the CPU oracle and FP64.

Bar (per the path's stated tolerance): max |gpu - oracle| <= 1e-5 of full
scale (1.0) and output SNR >= 120 dB against the oracle; every case also
checks the error against the FP64 model of the same schedule.
"""
import numpy as np
import pytest

import oracle_lib as O
import tan_amd as T
from helpers import crossfade_model, fft_conv_f64, snr_db

pytestmark = pytest.mark.gpu

UNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
BLOCKING = T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY
MAX_ABS = 1e-5
MIN_SNR = 120.0


def _irs(C, L, update, seed=1234):
    return T.signals.ir_bank(C, L, update=update, base_seed=seed)


def run_pair(ctx, L, B, C, nb, updates, method=UNIFORM, flags_at=None, finalize=False,
             x=None, oracle_channels=None):
    """Runs the GPU object and the oracle side by side through the same
    calls; updates maps block index -> IR bank (C, len)."""
    print(f"run_pair L={L} B={B} C={C} nb={nb} method={method:#x}", flush=True)
    conv = T.TANConvolution(ctx)
    assert conv.InitGpu(method, L, B, C) == T.AMF_OK
    orc = O.OracleConvolution(L, B, C, finalize_mode=finalize)
    if x is None:
        x = T.signals.input_bank(C, nb * B)
    y = np.zeros((C, nb * B), np.float32)
    yo = np.zeros((C, nb * B), np.float32)
    xfade_gpu, xfade_orc = [], []
    for b in range(nb):
        if b in updates:
            ir = updates[b]
            assert conv.UpdateResponseTD(ir, ir.shape[1], None, BLOCKING) == T.AMF_OK
            assert orc.update_td(ir) == O.AMF_OK
        fl = flags_at(b) if flags_at else None
        blk = np.ascontiguousarray(x[:, b * B:(b + 1) * B])
        out = np.zeros((C, B), np.float32)
        rc, done = conv.Process(blk, out, B, fl)
        assert rc == T.AMF_OK and done == B, (b, rc, done)
        rco, out_o, _ = orc.process(blk, flags=fl)
        assert rco == O.AMF_OK
        if finalize:
            assert conv.ProcessFinalize() == T.AMF_OK
            orc.finalize()
        y[:, b * B:(b + 1) * B] = out
        yo[:, b * B:(b + 1) * B] = out_o
        if conv.LastWasCrossfade():
            xfade_gpu.append(b)
        if orc.last_was_crossfade():
            xfade_orc.append(b)
    assert xfade_gpu == xfade_orc
    conv.Terminate()
    return x, y, yo, xfade_gpu


def assert_parity(y, yo, label=""):
    err = np.abs(y.astype(np.float64) - yo).max()
    snr = min(snr_db(yo[c], y[c]) for c in range(y.shape[0]) if np.any(yo[c]))
    print(f"{label} max|gpu-oracle|={err:.3e} snr={snr:.1f} dB", flush=True)
    assert err <= MAX_ABS
    assert snr >= MIN_SNR


def test_c1_cpu_config_parity(gpu_ctx):
    """C1: 2 ch, B=1024, L=16384 decaying noise (unit L2), 200 blocks, FP32;
    compared with the oracle and with FP64 direct convolution."""
    L, B, C, nb = 16384, 1024, 2, 200
    ir = _irs(C, L, 0)
    x, y, yo, xf = run_pair(gpu_ctx, L, B, C, nb, {0: ir})
    assert xf == [2]
    assert_parity(y, yo, "C1")
    for c in range(C):
        ref = crossfade_model(x[c], [ir[c]], [2], B, conv=lambda a, h, n: O.direct_conv_f64(a, h, n))
        err = np.abs(y[c] - ref).max()
        snr = snr_db(ref, y[c])
        print(f"C1 ch{c} vs FP64 direct: max abs {err:.3e}, SNR {snr:.1f} dB", flush=True)
        assert err <= MAX_ABS and snr >= MIN_SNR


def test_c2_shape_updates_and_crossfades(gpu_ctx):
    """B=256, L=65536 (P=256) with a fresh IR every 50 blocks; GPU vs oracle
    everywhere and vs the FP64 crossfade model."""
    L, B, C, nb = 65536, 256, 4, 320
    ups = {b: _irs(C, L, b // 50) for b in range(0, nb, 50)}
    x, y, yo, xf = run_pair(gpu_ctx, L, B, C, nb, ups)
    assert xf == [b + 2 for b in sorted(ups)]
    assert_parity(y, yo, "C2-shape")
    irs = [ups[b] for b in sorted(ups)]
    for c in (0, C - 1):
        ref = crossfade_model(x[c], [i[c] for i in irs], xf, B)
        err = np.abs(y[c] - ref).max()
        snr = snr_db(ref, y[c])
        print(f"C2-shape ch{c} vs FP64: max abs {err:.3e}, SNR {snr:.1f} dB", flush=True)
        assert err <= MAX_ABS and snr >= MIN_SNR


@pytest.mark.parametrize("B,L", [(64, 1000), (128, 4096), (512, 8192), (1024, 3000), (256, 256)])
def test_block_sizes(gpu_ctx, B, L):
    C, nb = 3, 40
    ups = {0: _irs(C, L, 0, seed=7), 17: _irs(C, L, 1, seed=7)}
    x, y, yo, xf = run_pair(gpu_ctx, L, B, C, nb, ups)
    assert xf == [2, 19]
    assert_parity(y, yo, f"B={B} L={L}")


def test_short_ir_update_zero_pads(gpu_ctx):
    """An IR shorter than Init's length is zero padded (TrueAudioNext.h:151-153)."""
    L, B, C, nb = 4096, 128, 2, 30
    short = _irs(C, 1000, 0)
    x, y, yo, xf = run_pair(gpu_ctx, L, B, C, nb, {0: short, 9: _irs(C, 2500, 1)})
    assert_parity(y, yo, "short IR")


def test_finalize_mode(gpu_ctx):
    L, B, C, nb = 8192, 256, 2, 60
    ups = {0: _irs(C, L, 0), 21: _irs(C, L, 1)}
    method = UNIFORM | T.TAN_CONVOLUTION_METHOD_USE_PROCESS_FINALIZE
    x, y, yo, xf = run_pair(gpu_ctx, L, B, C, nb, ups, method=method, finalize=True)
    assert xf == [2, 23]
    assert_parity(y, yo, "finalize")
    _, y2, _, _ = run_pair(gpu_ctx, L, B, C, nb, ups, x=x)
    np.testing.assert_allclose(y, y2, atol=1e-6)


def test_channel_flags(gpu_ctx):
    """STOP_INPUT drain (m_length samples), FLUSH_STREAM, untouched outputs of
    stopped channels (ConvolutionImpl.cpp:2868-2931)."""
    L, B, C, nb = 2048, 128, 4, 64
    P = L // B

    def flags_at(b):
        f = np.zeros(C, np.uint32)
        if 10 <= b < 10 + P + 6:
            f[1] = T.TAN_CONVOLUTION_CHANNEL_FLAG_STOP_INPUT
        if b in (20, 21):
            f[2] = T.TAN_CONVOLUTION_CHANNEL_FLAG_FLUSH_STREAM
        if 40 <= b < 44:
            f[3] = T.TAN_CONVOLUTION_CHANNEL_FLAG_STOP_INPUT | T.TAN_CONVOLUTION_CHANNEL_FLAG_FLUSH_STREAM
        return f

    ups = {0: _irs(C, L, 0), 30: _irs(C, L, 1)}
    x, y, yo, xf = run_pair(gpu_ctx, L, B, C, nb, ups, flags_at=flags_at)
    assert_parity(y, yo, "flags")


def test_next_free_channel_and_errors(gpu_ctx):
    L, B, C = 1024, 64, 3
    conv = T.TANConvolution(gpu_ctx)
    assert conv.InitGpu(UNIFORM, L, 96, C) == T.AMF_INVALID_ARG  # not a power of two
    assert conv.InitGpu(42, L, B, C) == T.AMF_NOT_SUPPORTED  # no such method
    assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
    assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_ALREADY_INITIALIZED
    assert conv.UpdateResponseTD(np.zeros((C, L + 1), np.float32)) == T.AMF_INVALID_ARG
    assert conv.UpdateResponseTD(_irs(C, L, 0), L, None, BLOCKING) == T.AMF_OK
    x = np.ones((C, B), np.float32)
    out = np.zeros((C, B), np.float32)
    rc, done = conv.Process(x[:, :32], out, 32)
    assert rc == T.AMF_OK and done == 0
    assert conv.Process(x, out, B)[0] == T.AMF_OK  # arms the drain counters
    stop = np.array([0, 1, 0], np.uint32)
    assert conv.GetNextFreeChannel(stop)[0] == T.AMF_NOT_FOUND
    for _ in range(L // B):
        assert conv.Process(x, out, B, stop)[0] == T.AMF_OK
    rc, idx = conv.GetNextFreeChannel(stop)
    assert rc == T.AMF_OK and idx == 1
    assert conv.GetNextFreeChannel(None)[0] == T.AMF_NOT_FOUND
    sentinel = np.full((C, B), 5.0, np.float32)
    assert conv.Process(x, sentinel, B, stop)[0] == T.AMF_OK
    assert np.all(sentinel[1] == 5.0)
    rc, _ = conv.Process(x, out, B, np.array([3, 3, 3], np.uint32))
    assert rc == T.AMF_WRONG_STATE
    assert conv.UpdateResponseFD(None) == T.AMF_NOT_SUPPORTED
    conv.Terminate()
    rc, ctx2 = T.TANCreateContext(version=(2 << 48))
    assert rc == T.AMF_TAN_UNSUPPORTED_VERSION and ctx2 is None


def test_update_while_crossfade_armed(gpu_ctx):
    """An update staged after the switch is armed is handed off after the
    crossfade (the reference would block the caller until then)."""
    L, B, C, nb = 2048, 128, 2, 30
    ups = {0: _irs(C, L, 0), 5: _irs(C, L, 1), 7: _irs(C, L, 2)}
    x, y, yo, xf = run_pair(gpu_ctx, L, B, C, nb, ups)
    assert_parity(y, yo, "armed update")


def test_partial_channel_update(gpu_ctx):
    """STOP_INPUT in UpdateResponseTD's flags keeps that channel's IR."""
    L, B, C, nb = 2048, 128, 3, 30
    conv = T.TANConvolution(gpu_ctx)
    conv.InitGpu(UNIFORM, L, B, C)
    orc = O.OracleConvolution(L, B, C)
    x = T.signals.input_bank(C, nb * B)
    skip = np.array([0, 1, 0], np.uint32)
    for b in range(nb):
        if b == 0:
            conv.UpdateResponseTD(_irs(C, L, 0), L, None, BLOCKING)
            orc.update_td(_irs(C, L, 0))
        if b == 12:
            conv.UpdateResponseTD(_irs(C, L, 1), L, skip, BLOCKING)
            orc.update_td(_irs(C, L, 1), flags=skip)
        blk = np.ascontiguousarray(x[:, b * B:(b + 1) * B])
        out = np.zeros((C, B), np.float32)
        conv.Process(blk, out)
        _, oo, _ = orc.process(blk)
        assert np.abs(out - oo).max() <= MAX_ABS, b
    conv.Terminate()


@pytest.mark.parametrize("chunks", [1, 3, 16])
def test_tail_chunking_invariant(gpu_ctx, chunks):
    L, B, C, nb = 16384, 256, 2, 40
    ir = _irs(C, L, 0)
    x = T.signals.input_bank(C, nb * B)
    conv = T.TANConvolution(gpu_ctx)
    conv.InitGpu(UNIFORM, L, B, C)
    conv.SetTailChunks(chunks)
    conv.UpdateResponseTD(ir, L, None, BLOCKING)
    orc = O.OracleConvolution(L, B, C)
    orc.update_td(ir)
    for b in range(nb):
        blk = np.ascontiguousarray(x[:, b * B:(b + 1) * B])
        out = np.zeros((C, B), np.float32)
        conv.Process(blk, out)
        _, oo, _ = orc.process(blk)
        assert np.abs(out - oo).max() <= MAX_ABS
    conv.Terminate()


@pytest.mark.parametrize("fused", [True, False])
def test_device_memory_variants(gpu_ctx, fused):
    """Process/UpdateResponseTD on device buffers (the cl_mem stand-ins)
    equal the host-memory path bit for bit, with the head and the next
    block's tail in one launch (fused) or in two."""
    import torch
    L, B, C, nb = 8192, 256, 3, 24
    ir = _irs(C, L, 0)
    x = T.signals.input_bank(C, nb * B)
    conv_h = T.TANConvolution(gpu_ctx)
    conv_d = T.TANConvolution(gpu_ctx)
    conv_h.InitGpu(UNIFORM, L, B, C)
    conv_d.InitGpu(UNIFORM, L, B, C)
    conv_d.SetFused(fused)
    conv_h.UpdateResponseTD(ir, L, None, BLOCKING)
    dir_ = torch.from_numpy(ir).cuda()
    torch.cuda.synchronize()
    assert conv_d.UpdateResponseTD(dir_, L, None, BLOCKING) == T.AMF_OK
    dx = torch.from_numpy(x).cuda()
    dy = torch.zeros_like(dx)
    torch.cuda.synchronize()
    yh = np.zeros_like(x)
    for b in range(nb):
        out = np.zeros((C, B), np.float32)
        conv_h.Process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]), out)
        yh[:, b * B:(b + 1) * B] = out
        rc, done = conv_d.Process(dx[:, b * B:], dy[:, b * B:], B)
        assert rc == T.AMF_OK and done == B
    conv_d.Synchronize()
    np.testing.assert_array_equal(dy.cpu().numpy(), yh)
    conv_h.Terminate()
    conv_d.Terminate()


def test_multiblock_device_process_with_updates(gpu_ctx):
    """One device Process call over many blocks equals block-by-block calls,
    including a crossfade inside the call: bit for bit with a tail every
    block, up to rounding with multi-block passes (a call's heads run
    together, with the pass for the blocks after them), and the oracle's."""
    import torch
    L, B, C, nb = 4096, 128, 2, 40
    x = T.signals.input_bank(C, nb * B)
    dx = torch.from_numpy(x).cuda()
    orc = O.OracleConvolution(L, B, C)
    yo = np.zeros_like(x)
    for b in range(nb):
        if b in (0, 20):
            orc.update_td(_irs(C, L, b))
        _, o, _ = orc.process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]))
        yo[:, b * B:(b + 1) * B] = o
    for kp in (1, 4):
        outs = []
        for per_call in (1, 20):
            conv = T.TANConvolution(gpu_ctx, pass_blocks=kp)
            conv.InitGpu(UNIFORM, L, B, C)
            dy = torch.zeros_like(dx)
            torch.cuda.synchronize()
            for b0 in range(0, nb, per_call):
                if b0 in (0, 20):
                    conv.UpdateResponseTD(_irs(C, L, b0), L, None, BLOCKING)
                rc, done = conv.Process(dx[:, b0 * B:], dy[:, b0 * B:], per_call * B)
                assert rc == T.AMF_OK and done == per_call * B
            conv.Synchronize()
            outs.append(dy.cpu().numpy())
            conv.Terminate()
        if kp == 1:
            np.testing.assert_array_equal(outs[0], outs[1])
        else:
            assert np.abs(outs[0].astype(np.float64) - outs[1]).max() <= 1e-6
        for y in outs:
            assert_parity(y, yo, f"multiblock kp={kp}")


def test_autotune_keeps_state_and_parity(gpu_ctx):
    """The autotuner's dry runs leave a fresh object unchanged; it refuses
    once blocks have been processed."""
    L, B, C, nb = 65536, 256, 8, 12
    conv = T.TANConvolution(gpu_ctx)
    conv.InitGpu(UNIFORM, L, B, C)
    ir = _irs(C, L, 0)
    conv.UpdateResponseTD(ir, L, None, BLOCKING)
    rc, chunks, us = conv.Autotune()
    assert rc == T.AMF_OK and chunks >= 1 and us > 0
    orc = O.OracleConvolution(L, B, C)
    orc.update_td(ir)
    x = T.signals.input_bank(C, nb * B)
    for b in range(nb):
        blk = np.ascontiguousarray(x[:, b * B:(b + 1) * B])
        out = np.zeros((C, B), np.float32)
        conv.Process(blk, out)
        _, oo, _ = orc.process(blk)
        assert np.abs(out - oo).max() <= MAX_ABS
    assert conv.Autotune()[0] == T.AMF_WRONG_STATE
    conv.Terminate()


def test_graph_replay_matches_direct_launches(gpu_ctx):
    """Device Process over many blocks replays a captured graph of tick
    launches; results equal direct launches bit for bit, across batch
    boundaries (64 blocks) and crossfades."""
    import torch
    L, B, C, nb = 16384, 256, 4, 150
    x = T.signals.input_bank(C, nb * B)
    dx = torch.from_numpy(x).cuda()
    outs = []
    for graph in (True, False):
        conv = T.TANConvolution(gpu_ctx)
        conv.InitGpu(UNIFORM, L, B, C)
        conv.SetGraph(graph)
        dy = torch.zeros_like(dx)
        torch.cuda.synchronize()
        for b0, n in ((0, 70), (70, 5), (75, 75)):
            conv.UpdateResponseTD(_irs(C, L, b0), L, None, BLOCKING)
            rc, done = conv.Process(dx[:, b0 * B:], dy[:, b0 * B:], n * B)
            assert rc == T.AMF_OK and done == n * B
        conv.Synchronize()
        outs.append(dy.cpu().numpy())
        conv.Terminate()
    np.testing.assert_array_equal(outs[0], outs[1])


def test_objects_keep_their_context():
    """A context terminated while objects made on it still exist stays alive
    until the last of them goes (the objects hold references)."""
    import tan_amd as TT
    ctx = TT.TANContext(device=0)
    L, B, C = 4096, 256, 2
    conv = T.TANConvolution(ctx)
    assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
    assert conv.UpdateResponseTD(_irs(C, L, 3), L, None, BLOCKING) == T.AMF_OK
    assert ctx.Terminate() == T.AMF_OK
    x = T.signals.input_bank(C, 4 * B)
    for b in range(4):
        out = np.zeros((C, B), np.float32)
        rc, done = conv.Process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]), out)
        assert rc == T.AMF_OK and done == B
    assert np.isfinite(out).all()
    assert conv.Terminate() == T.AMF_OK


def test_context_on_caller_streams(gpu_ctx):
    """A context on the caller's streams (InitOpenCL with the application's
    queues): Process and updates run on them, work the caller queued on the
    same stream needs no ordering, and the output is the oracle's; the
    context drains but keeps the caller's stream."""
    import torch
    L, B, C, nb = 8192, 256, 2, 30
    s = torch.cuda.Stream()
    ctx = T.TANContext(device=0, conv_stream=s, general_stream=s)
    lib = ctx._lib
    assert lib.tanx_ctx_process_stream(ctx.handle) == s.cuda_stream
    assert lib.tanx_ctx_update_stream(ctx.handle) == s.cuda_stream
    ir = _irs(C, L, 5)
    x = T.signals.input_bank(C, nb * B, base_seed=21)
    with torch.cuda.stream(s):
        irs = torch.from_numpy(ir).cuda()
        xd = torch.from_numpy(x).cuda()
        yd = torch.zeros_like(xd)
        conv = T.TANConvolution(ctx)
        assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
        assert conv.UpdateResponseTD(irs, L, None, BLOCKING) == T.AMF_OK
        for b0, n in ((0, 7), (7, 1), (8, 22)):
            rc, done = conv.Process(xd[:, b0 * B:], yd[:, b0 * B:], n * B)
            assert rc == T.AMF_OK and done == n * B
        y = yd.cpu().numpy()  # ordered on s
    conv.Terminate()
    ctx.Terminate()
    s.synchronize()  # still a valid stream
    orc = O.OracleConvolution(L, B, C)
    orc.update_td(ir)
    yo = np.zeros_like(x, dtype=np.float64)
    for b in range(nb):
        _, o, _ = orc.process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]))
        yo[:, b * B:(b + 1) * B] = o
    assert_parity(y, yo, "caller streams")
    # no place for the handle
    assert lib.tanx_ctx_create_on_streams(T.TAN_FULL_VERSION, 0, None, None, None) != T.AMF_OK
