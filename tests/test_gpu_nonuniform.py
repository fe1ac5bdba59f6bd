"""Non-uniform partitioned objects (TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM
and its aliases): the first frame of taps as uniform partitions of B, the
rest as frame partitions of M * B whose multiply-accumulate is spread over
the M blocks of a frame.  The output is the same convolution with the same
B-sample latency, so it is pinned to the FP64 model of the time-varying
convolution (the intended crossfade, at the non-uniform switch blocks) and to
the uniform engine on the same inputs."""
import numpy as np
import pytest

import tan_amd as T
from helpers import crossfade_model, fft_conv_f64, snr_db

pytestmark = pytest.mark.gpu
NU = T.TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM
UNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
BLOCKING = T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY


def nu_switch_block(update_before, M, last_switch=None):
    """Block of the crossfade for an update issued before block
    `update_before`: armed one block later (or, when issued while another
    switch is armed, at the block after that switch), then the first frame
    boundary at least M + 1 blocks after arming."""
    armed = update_before + 1
    if last_switch is not None and armed <= last_switch:
        armed = last_switch + 1
    return -(-(armed + M + 1) // M) * M


def _run(ctx, method, L, B, C, x, irs, updates, multiple=0, precision=T.TAN_IR_PRECISION_FP32,
         device=True):
    """Process x block by block; irs[k] is installed before block updates[k]."""
    import torch
    conv = T.TANConvolution(ctx, ir_precision=precision, nu_multiple=multiple)
    assert conv.InitGpu(method, L, B, C) == T.AMF_OK
    info = conv.Info()
    nb = x.shape[1] // B
    y = np.zeros_like(x)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    k = 0
    for b in range(nb):
        while k < len(updates) and updates[k] == b:
            assert conv.UpdateResponseTD(irs[k], L, None, BLOCKING if b == 0 else 0) == T.AMF_OK
            k += 1
        if device:
            rc, done = conv.Process(xd[:, b * B:], yd[:, b * B:], B)
        else:
            rc, done = conv.Process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]), y[:, b * B:(b + 1) * B], B)
        assert rc == T.AMF_OK and done == B
    conv.Synchronize()
    if device:
        y = yd.cpu().numpy()
    conv.Terminate()
    return y, info


@pytest.mark.parametrize("L,B,C,M", [(16384, 256, 2, 4), (65536, 128, 3, 8), (8192, 64, 1, 2),
                                     (131072, 256, 2, 16), (40000, 512, 2, 4), (20000, 1024, 1, 2)])
def test_static_ir_matches_fp64(gpu_ctx, L, B, C, M):
    nb = 6 * M + 40
    x = T.signals.input_bank(C, nb * B, base_seed=70)
    ir = T.signals.ir_bank(C, L, 0, base_seed=71)
    y, info = _run(gpu_ctx, NU, L, B, C, x, [ir], [0], multiple=M)
    assert info["nu_multiple"] == M
    s = nu_switch_block(0, M)
    for c in range(C):
        ref = crossfade_model(x[c], [ir[c]], [s], B)
        assert np.abs(y[c] - ref).max() <= 1e-5, f"channel {c}"
        assert snr_db(ref[s * B:], y[c, s * B:]) >= 120.0
        assert np.abs(y[c, :s * B]).max() == 0.0  # silent before the first switch


def test_updates_crossfade_at_frame_boundaries(gpu_ctx):
    L, B, C, M = 32768, 256, 2, 8
    nb = 120
    x = T.signals.input_bank(C, nb * B, base_seed=80)
    ups = [0, 19, 40, 43, 77]  # 43: while the switch of 40 is armed
    irs = [T.signals.ir_bank(C, L, k, base_seed=81) for k in range(len(ups))]
    y, _ = _run(gpu_ctx, NU, L, B, C, x, irs, ups, multiple=M)
    # an update while one is armed switches at the first boundary after the
    # armed switch, once it has been handed off and armed in turn
    sw = []
    for u in ups:
        sw.append(nu_switch_block(u, M, sw[-1] if sw else None))
    for c in range(C):
        ref = crossfade_model(x[c], [h[c] for h in irs], sw, B)
        assert np.abs(y[c] - ref).max() <= 1e-5, f"channel {c}: switches {sw}"
        assert snr_db(ref, y[c]) >= 120.0


def test_nonuniform_equals_uniform_after_fade_in(gpu_ctx):
    L, B, C = 131072, 256, 4
    nb = 200
    x = T.signals.input_bank(C, nb * B, base_seed=90)
    ir = T.signals.ir_bank(C, L, 0, base_seed=91)
    yn, info = _run(gpu_ctx, NU, L, B, C, x, [ir], [0])
    yu, _ = _run(gpu_ctx, UNIFORM, L, B, C, x, [ir], [0])
    M = info["nu_multiple"]
    assert M > 1, "the library picks frames for a long IR"
    s = nu_switch_block(0, M)
    # block s fades the IR in; from block s + 1 both are the plain convolution
    assert np.abs(yn[:, (s + 1) * B:] - yu[:, (s + 1) * B:]).max() <= 1e-5


def test_host_memory_path_and_fp16(gpu_ctx):
    L, B, C, M = 16384, 128, 2, 4
    nb = 60
    x = T.signals.input_bank(C, nb * B, base_seed=95)
    ir = T.signals.ir_bank(C, L, 0, base_seed=96)
    yd, _ = _run(gpu_ctx, NU, L, B, C, x, [ir], [0], multiple=M)
    yh, _ = _run(gpu_ctx, NU, L, B, C, x, [ir], [0], multiple=M, device=False)
    np.testing.assert_array_equal(yd, yh)
    y16, _ = _run(gpu_ctx, NU, L, B, C, x, [ir], [0], multiple=M, precision=T.TAN_IR_PRECISION_FP16)
    s = nu_switch_block(0, M)
    for c in range(C):
        ref = crossfade_model(x[c], [ir[c]], [s], B)
        assert snr_db(ref[s * B:], y16[c, s * B:]) >= 60.0


def test_flush_and_stop_flags(gpu_ctx):
    """A flushed channel restarts from silence as in the uniform engine."""
    import torch
    L, B, C, M = 16384, 256, 2, 4
    nb = 64
    x = T.signals.input_bank(C, nb * B, base_seed=97)
    ir = T.signals.ir_bank(C, L, 0, base_seed=98)
    outs = []
    for method, mult in ((NU, M), (UNIFORM, 0)):
        conv = T.TANConvolution(gpu_ctx, nu_multiple=mult)
        assert conv.InitGpu(method, L, B, C) == T.AMF_OK
        assert conv.UpdateResponseTD(ir, L, None, BLOCKING) == T.AMF_OK
        xd = torch.from_numpy(x).cuda()
        yd = torch.zeros_like(xd)
        for b in range(nb):
            flags = None
            if b == 30:
                flags = np.array([T.TAN_CONVOLUTION_CHANNEL_FLAG_FLUSH_STREAM, 0], np.uint32)
            rc, done = conv.Process(xd[:, b * B:], yd[:, b * B:], B, flags)
            assert rc == T.AMF_OK and done == B
        conv.Synchronize()
        outs.append(yd.cpu().numpy())
        conv.Terminate()
    s = nu_switch_block(0, M)
    np.testing.assert_allclose(outs[0][:, (s + 1) * B:], outs[1][:, (s + 1) * B:], atol=1e-5, rtol=0)
    # channel 0 after the flush: the convolution of the input from block 30 on
    ref = fft_conv_f64(x[0, 30 * B:], ir[0])
    assert np.abs(outs[0][0, 30 * B:] - ref).max() <= 1e-5


def test_group_launch_with_nonuniform_objects(gpu_ctx):
    import torch
    specs = [(NU, 65536, 256, 2, 8), (UNIFORM, 8192, 256, 1, 0), (NU, 20000, 128, 3, 4),
             (NU, 262144, 512, 1, 0), (NU, 4096, 64, 2, 2)]
    base, nticks = 64, 160
    objs, alone, xs, yg, ya = [], [], [], [], []
    for i, (meth, L, B, C, mult) in enumerate(specs):
        nbk = nticks * base // B
        x = T.signals.input_bank(C, nbk * B, base_seed=110 + i)
        ir = T.signals.ir_bank(C, L, 0, base_seed=120 + i)
        for lst in (objs, alone):
            c = T.TANConvolution(gpu_ctx, nu_multiple=mult)
            assert c.InitGpu(meth, L, B, C) == T.AMF_OK
            assert c.UpdateResponseTD(ir, L, None, BLOCKING) == T.AMF_OK
            lst.append(c)
        xs.append(torch.from_numpy(x).cuda())
        yg.append(torch.zeros(C, nbk * B, device="cuda"))
        ya.append(torch.zeros(C, nbk * B, device="cuda"))
    for t in range(nticks):
        due = [i for i, s in enumerate(specs) if (t * base) % s[2] == 0]
        blk = {i: (t * base) // specs[i][2] for i in due}
        rc = T.GroupProcess([objs[i] for i in due], [xs[i][:, blk[i] * specs[i][2]:] for i in due],
                            [yg[i][:, blk[i] * specs[i][2]:] for i in due])
        assert rc == T.AMF_OK
        for i in due:
            B = specs[i][2]
            rc, done = alone[i].Process(xs[i][:, blk[i] * B:], ya[i][:, blk[i] * B:], B)
            assert rc == T.AMF_OK and done == B
    for c in objs + alone:
        c.Synchronize()
    for i in range(len(specs)):
        np.testing.assert_array_equal(yg[i].cpu().numpy(), ya[i].cpu().numpy(), err_msg=f"object {i}")
    for c in objs + alone:
        c.Terminate()
