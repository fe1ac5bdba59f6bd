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


def nu_switch_block(update_before, M, last_switch=None, phase=0):
    """Block of the crossfade for an update issued before block
    `update_before`: armed one block later (or, when issued while another
    switch is armed, at the block after that switch; issued before the
    crossfade block itself, a block later still), then the first frame
    boundary at least M + 1 blocks after arming.  Frame boundaries lie where
    (block + phase) % M == 0."""
    handed = update_before
    if last_switch is not None and update_before == last_switch:
        handed += 1
    armed = handed + 1
    if last_switch is not None and armed <= last_switch:
        armed = last_switch + 1
    return -(-(armed + phase + M + 1) // M) * M - phase


def nu_schedule(updates, M, phase, nb):
    """Crossfades of a non-uniform object as (block, update index) pairs: a
    block-level model of the switch state machine.  An update issued before
    block b is handed off at the end of the first block >= b that does not
    crossfade and armed at the next block with no switch armed; an armed
    switch falls on the first frame boundary ((block + phase) % M == 0) at
    least M + 1 blocks after arming.  An update issued while a switch is
    armed is held for the next one, and an update that has not been armed
    yet is replaced by a later one."""
    pending = ready = armed = False
    waiting = held = None  # update in the slot to arm next / held while armed
    switch_at = cur = None
    out = []
    k = 0
    for b in range(nb):
        while k < len(updates) and updates[k] == b:
            if armed:
                held = k
            else:
                waiting = k
            pending = True
            k += 1
        xfade = False
        if not pending and ready and not armed:
            ready, armed, cur = False, True, waiting
            switch_at = -(-(b + phase + M + 1) // M) * M - phase
        elif armed and b == switch_at:
            xfade = True
            out.append((b, cur))
            armed = False
            waiting, held = held, None
        if not xfade and pending:
            pending, ready = False, True
    return out


def _run(ctx, method, L, B, C, x, irs, updates, multiple=0, precision=T.TAN_IR_PRECISION_FP32,
         device=True, phase=0):
    """Process x block by block; irs[k] is installed before block updates[k]."""
    import torch
    conv = T.TANConvolution(ctx, ir_precision=precision, nu_multiple=multiple, nu_phase=phase)
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


@pytest.mark.parametrize("phase", [1, 3, 6, 7, 13])
def test_frame_phase_shifts_frames_and_switches(gpu_ctx, phase):
    """A frame phase moves the frame boundaries (and with them the switch
    blocks) by `phase` blocks, modulo M; the output is the same time-varying
    convolution at those switch blocks.  Updates 9 and 17 fall on or next to
    the switch blocks of some phases."""
    L, B, C, M = 32768, 128, 2, 8
    nb = 100
    x = T.signals.input_bank(C, nb * B, base_seed=130)
    ups = [0, 9, 17, 40, 58]
    irs = [T.signals.ir_bank(C, L, k, base_seed=131) for k in range(len(ups))]
    y, info = _run(gpu_ctx, NU, L, B, C, x, irs, ups, multiple=M, phase=phase)
    assert info["nu_multiple"] == M
    sched = nu_schedule(ups, M, phase % M, nb)
    sw = [s for s, _ in sched]
    assert all((s + phase) % M == 0 for s in sw)
    for c in range(C):
        ref = crossfade_model(x[c], [irs[k][c] for _, k in sched], sw, B)
        assert np.abs(y[c] - ref).max() <= 1e-5, f"channel {c}: switches {sw}"
        assert snr_db(ref, y[c]) >= 120.0


def test_phased_objects_in_one_group(gpu_ctx):
    """M objects with phases 0..M-1 advanced in one group launch: each tick
    runs the frame kernel of one object, and every object's output equals
    the same object advanced alone."""
    import torch
    L, B, C, M = 65536, 256, 3, 4
    nb = 40
    objs, alone, xs, yg, ya = [], [], [], [], []
    for p in range(M):
        x = T.signals.input_bank(C, nb * B, base_seed=140 + p)
        ir = T.signals.ir_bank(C, L, 0, base_seed=150 + p)
        for lst in (objs, alone):
            c = T.TANConvolution(gpu_ctx, nu_multiple=M, nu_phase=p)
            assert c.InitGpu(NU, L, B, C) == T.AMF_OK
            assert c.UpdateResponseTD(ir, L, None, BLOCKING) == T.AMF_OK
            lst.append(c)
        xs.append(torch.from_numpy(x).cuda())
        yg.append(torch.zeros(C, nb * B, device="cuda"))
        ya.append(torch.zeros(C, nb * B, device="cuda"))
    for b in range(nb):
        assert T.GroupProcess(objs, [xi[:, b * B:] for xi in xs], [yi[:, b * B:] for yi in yg]) == T.AMF_OK
        for p in range(M):
            rc, done = alone[p].Process(xs[p][:, b * B:], ya[p][:, b * B:], B)
            assert rc == T.AMF_OK and done == B
    for c in objs + alone:
        c.Synchronize()
    for p in range(M):
        np.testing.assert_array_equal(yg[p].cpu().numpy(), ya[p].cpu().numpy(), err_msg=f"phase {p}")
        s = nu_switch_block(0, M, phase=p)
        ref = fft_conv_f64(xs[p].cpu().numpy()[0], T.signals.ir_bank(C, L, 0, base_seed=150 + p)[0])
        assert np.abs(yg[p].cpu().numpy()[0, (s + 1) * B:] - ref[(s + 1) * B:]).max() <= 1e-5
    for c in objs + alone:
        c.Terminate()
