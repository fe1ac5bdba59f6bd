"""Multi-block tail passes (tanx_conv_options.pass_blocks): a launch runs
up to K heads and, when the next block has no partial sums yet, a pass that
computes the next K blocks' sums from the rows before its heads; each later
head adds the partitions since that launch, and a switch armed at a launch's
last block makes its pass compute the new filter's sums.  The output is the
reference's schedule: equal to the oracle (and the FP64 model) for every K,
under dense updates that arm switches at every offset, updates while a
switch is armed, channel flushes, the host-memory path and calls of any
number of blocks."""
import numpy as np
import pytest

import oracle_lib as O
import tan_amd as T
from helpers import crossfade_model, snr_db

pytestmark = pytest.mark.gpu

UNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
BLOCKING = T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY


def _run(ctx, L, B, C, nb, updates, K, device=True, flags_at=None, seed=500):
    import torch
    print(f"passes L={L} B={B} C={C} K={K} device={device}", flush=True)
    conv = T.TANConvolution(ctx, pass_blocks=K)
    assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
    assert conv.Info()["pass_blocks"] == K
    orc = O.OracleConvolution(L, B, C)
    x = T.signals.input_bank(C, nb * B, base_seed=seed)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    y = np.zeros_like(x)
    yo = np.zeros_like(x)
    xf_gpu, xf_orc = [], []
    for b in range(nb):
        if b in updates:
            ir = updates[b]
            assert conv.UpdateResponseTD(ir, L, None, BLOCKING) == T.AMF_OK
            assert orc.update_td(ir) == O.AMF_OK
        fl = flags_at(b) if flags_at else None
        blk = np.ascontiguousarray(x[:, b * B:(b + 1) * B])
        if device:
            rc, done = conv.Process(xd[:, b * B:], yd[:, b * B:], B, fl)
        else:
            out = np.zeros((C, B), np.float32)
            rc, done = conv.Process(blk, out, B, fl)
            y[:, b * B:(b + 1) * B] = out
        assert rc == T.AMF_OK and done == B
        rco, out_o, _ = orc.process(blk, flags=fl)
        yo[:, b * B:(b + 1) * B] = out_o
        if conv.LastWasCrossfade():
            xf_gpu.append(b)
        if orc.last_was_crossfade():
            xf_orc.append(b)
    conv.Synchronize()
    if device:
        y = yd.cpu().numpy()
    conv.Terminate()
    assert xf_gpu == xf_orc
    return x, y, yo, xf_gpu


def uniform_schedule(updates, nb):
    """Crossfades as (block, update index): an update issued before block b
    is handed off at the end of the first block >= b that does not
    crossfade, armed at the next block with no switch armed, and crossfades
    one block later; an update issued while a switch is armed is held for
    the next one, and one not yet armed is replaced by a later one."""
    pending = ready = armed = False
    waiting = held = cur = None
    switch_at = None
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
            switch_at = b + 1
        elif armed and b == switch_at:
            xfade = True
            out.append((b, cur))
            armed = False
            waiting, held = held, None
        if not xfade and pending:
            pending, ready = False, True
    return out


def _updates(C, L, blocks, seed=600):
    return {b: T.signals.ir_bank(C, L, update=i, base_seed=seed) for i, b in enumerate(blocks)}


DENSE = [0, 7, 8, 20, 21, 22, 40, 44, 45, 61, 62, 63, 64, 90]


@pytest.mark.parametrize("K", [1, 2, 3, 4, 5, 8, 11, 12, 16])
def test_dense_updates_every_pass_length(gpu_ctx, K):
    L, B, C, nb = 8192, 256, 3, 110
    ups = _updates(C, L, DENSE)
    x, y, yo, xf = _run(gpu_ctx, L, B, C, nb, ups, K)
    err = np.abs(y.astype(np.float64) - yo).max()
    assert err <= 1e-5, err
    for c in range(C):
        assert snr_db(yo[c], y[c]) >= 120.0
    # and the FP64 model of the reference's schedule
    order = sorted(ups)
    sched = uniform_schedule(order, nb)
    assert [s for s, _ in sched] == xf
    for c in range(C):
        ref = crossfade_model(x[c], [ups[order[k]][c] for _, k in sched], xf, B)
        assert np.abs(y[c] - ref).max() <= 1e-5


@pytest.mark.parametrize("K", [2, 4])
def test_flushes_and_host_path(gpu_ctx, K):
    L, B, C, nb = 4096, 128, 2, 70

    def flags_at(b):
        if b in (13, 30, 31, 46):
            return np.array([T.TAN_CONVOLUTION_CHANNEL_FLAG_FLUSH_STREAM, 0], np.uint32)
        return None

    ups = _updates(C, L, [0, 17, 33, 34, 50])
    for device in (True, False):
        x, y, yo, _ = _run(gpu_ctx, L, B, C, nb, ups, K, device=device, flags_at=flags_at, seed=501)
        err = np.abs(y.astype(np.float64) - yo).max()
        assert err <= 1e-5, (device, err)


def test_default_pass_length(gpu_ctx):
    """Uniform objects with 16 or more partitions use 4-block passes by
    default, at any channel count; shorter IRs and non-uniform objects keep a
    tail every block."""
    for L, B, C, want in ((65536, 256, 2, 4), (2048, 256, 2, 1), (16384, 1024, 2, 4), (65536, 256, 8192, 4),
                          (65536, 256, 4096, 4)):
        conv = T.TANConvolution(gpu_ctx)
        assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
        assert conv.Info()["pass_blocks"] == want, (L, B, C)
        conv.Terminate()
    conv = T.TANConvolution(gpu_ctx)
    assert conv.InitGpu(T.TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM, 65536, 256, 2) == T.AMF_OK
    assert conv.Info()["pass_blocks"] == 1
    conv.Terminate()
    conv = T.TANConvolution(gpu_ctx, pass_blocks=17)
    assert conv.InitGpu(UNIFORM, 65536, 256, 2) == T.AMF_INVALID_ARG


def _run_calls(ctx, L, B, C, calls, updates, K, precision=T.TAN_IR_PRECISION_FP32, seed=700):
    """Process in calls of calls[i] blocks (device memory), with
    updates[b] issued before the call that starts at block b."""
    import torch
    conv = T.TANConvolution(ctx, ir_precision=precision, pass_blocks=K)
    assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
    nb = sum(calls)
    x = T.signals.input_bank(C, nb * B, base_seed=seed)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    b = 0
    for n in calls:
        if b in updates:
            assert conv.UpdateResponseTD(updates[b], L, None, BLOCKING) == T.AMF_OK
        rc, done = conv.Process(xd[:, b * B:], yd[:, b * B:], n * B)
        assert rc == T.AMF_OK and done == n * B
        b += n
    conv.Synchronize()
    y = yd.cpu().numpy()
    conv.Terminate()
    return x, y


CALLS = [1, 3, 7, 2, 13, 5, 1, 1, 9, 4, 6, 11, 2, 8]


@pytest.mark.parametrize("K,B,precision", [(4, 256, T.TAN_IR_PRECISION_FP32), (3, 128, T.TAN_IR_PRECISION_FP32),
                                           (2, 512, T.TAN_IR_PRECISION_FP32), (4, 64, T.TAN_IR_PRECISION_FP32),
                                           (4, 256, T.TAN_IR_PRECISION_FP16), (8, 256, T.TAN_IR_PRECISION_FP32),
                                           (6, 128, T.TAN_IR_PRECISION_FP32), (8, 64, T.TAN_IR_PRECISION_BF16),
                                           (16, 256, T.TAN_IR_PRECISION_FP32), (13, 128, T.TAN_IR_PRECISION_FP16)])
def test_multi_block_calls_batch_heads(gpu_ctx, K, B, precision):
    """Calls of several blocks launch up to K heads together with the pass
    for the blocks after them (which reads the rows before the batch): the
    output is the oracle's, with updates (crossfades, catch-up passes)
    between calls, and that of one block per call up to rounding (the two
    split each block's sum between pass and head differently)."""
    L, C = 32 * B, 3
    nb = sum(CALLS)
    starts = np.cumsum([0] + CALLS[:-1])
    ups = _updates(C, L, [int(starts[i]) for i in (0, 2, 3, 5, 6, 7, 10, 12)], seed=601)
    print(f"batched heads K={K} B={B} precision={precision}", flush=True)
    x, y = _run_calls(gpu_ctx, L, B, C, CALLS, ups, K, precision)
    _, y1 = _run_calls(gpu_ctx, L, B, C, [1] * nb, ups, K, precision)
    assert np.abs(y.astype(np.float64) - y1).max() <= 1e-6
    if precision == T.TAN_IR_PRECISION_FP32:
        orc = O.OracleConvolution(L, B, C)
        yo = np.zeros_like(x)
        for b in range(nb):
            if b in ups:
                assert orc.update_td(ups[b]) == O.AMF_OK
            rco, out_o, _ = orc.process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]))
            yo[:, b * B:(b + 1) * B] = out_o
        err = np.abs(y.astype(np.float64) - yo).max()
        assert err <= 1e-5, err
        for c in range(C):
            assert snr_db(yo[c], y[c]) >= 120.0


def test_kernel_timing_counts_blocks(gpu_ctx):
    """With kernel timing on, a batched call is timed too: the timed
    launches (heads batched with the next blocks' pass) cover every block
    and count their algorithmic bytes."""
    import torch
    L, B, C, nb = 16384, 256, 4, 40
    conv = T.TANConvolution(gpu_ctx, pass_blocks=4)
    assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
    assert conv.UpdateResponseTD(T.signals.ir_bank(C, L, base_seed=9), L, None, BLOCKING) == T.AMF_OK
    x = torch.from_numpy(T.signals.input_bank(C, nb * B, base_seed=10)).cuda()
    y = torch.zeros_like(x)
    assert conv.EnableTiming(True) == T.AMF_OK
    conv.ReadTiming(reset=True)
    rc, done = conv.Process(x, y, nb * B)
    assert rc == T.AMF_OK and done == nb * B
    kt = conv.ReadTiming(reset=True)
    conv.Terminate()
    assert kt["blocks"] == nb
    assert kt["head_launches"] + kt["tail_launches"] < nb  # several heads share a launch
    assert kt["tail_ms"] > 0 and kt["tail_alg_bytes"] > 0


@pytest.mark.parametrize("chunks", [1, 3, 16, 30])
def test_pass_chunk_counts(gpu_ctx, chunks):
    """A pass splits a channel's partitions into chunks whose rows the
    heads sum, whatever their number: the oracle's output at every chunk
    count, with updates and uneven calls."""
    import torch
    L, B, C = 8192, 256, 3
    calls = [1, 5, 2, 9, 4, 1, 1, 7, 3, 8]
    nb = sum(calls)
    starts = np.cumsum([0] + calls[:-1])
    ups = _updates(C, L, [int(starts[i]) for i in (0, 2, 5, 7)], seed=602)
    conv = T.TANConvolution(gpu_ctx, pass_blocks=4)
    assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
    assert conv.SetTailChunks(chunks) == T.AMF_OK
    print(f"pass chunks={chunks}", flush=True)
    x = T.signals.input_bank(C, nb * B, base_seed=703)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    b = 0
    for n in calls:
        if b in ups:
            assert conv.UpdateResponseTD(ups[b], L, None, BLOCKING) == T.AMF_OK
        rc, done = conv.Process(xd[:, b * B:], yd[:, b * B:], n * B)
        assert rc == T.AMF_OK and done == n * B
        b += n
    conv.Synchronize()
    y = yd.cpu().numpy()
    conv.Terminate()
    orc = O.OracleConvolution(L, B, C)
    yo = np.zeros_like(x)
    for b in range(nb):
        if b in ups:
            assert orc.update_td(ups[b]) == O.AMF_OK
        _, out_o, _ = orc.process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]))
        yo[:, b * B:(b + 1) * B] = out_o
    assert np.abs(y.astype(np.float64) - yo).max() <= 1e-5


def _oracle_run(L, B, x, irs, nb):
    C = x.shape[0]
    orc = O.OracleConvolution(L, B, C)
    yo = np.zeros_like(x)
    for b in range(nb):
        if b in irs:
            assert orc.update_td(irs[b]) == O.AMF_OK
        yo[:, b * B:(b + 1) * B] = orc.process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]))[1]
    return yo


@pytest.mark.parametrize("K", [4, 8, 12, 16])
def test_pass_phases_alone_and_grouped(gpu_ctx, K):
    """Pass phases 0..K-1 (tanx_conv_options.pass_phase) of K-block passes:
    each object alone with calls of uneven size and dense updates, and the K
    together in one group launch per block, all equal to the oracle; in the
    group, a tick with no switch in flight streams exactly one object's
    pass."""
    import torch
    L, B, C, nb = 16384, 128, 2, 80
    x = T.signals.input_bank(C, nb * B, base_seed=41)
    ups = {0: 0, 9: 1, 10: 2, 23: 3, 24: 4, 25: 5, 50: 6}
    irs = {b: T.signals.ir_bank(C, L, k, base_seed=42) for b, k in ups.items()}
    yo = _oracle_run(L, B, x, irs, nb)
    xd = torch.from_numpy(x).cuda()
    sizes = [1, 2, 5, 3, 7, 1, 4]
    cuts = sorted(set(list(irs) + [nb]))
    for phase in range(K):
        conv = T.TANConvolution(gpu_ctx, pass_blocks=K, pass_phase=phase)
        assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
        yd = torch.zeros_like(xd)
        t = k = 0
        while t < nb:
            if t in irs:
                assert conv.UpdateResponseTD(torch.from_numpy(irs[t]).cuda(), L, None, BLOCKING) == T.AMF_OK
            n = min(sizes[k % len(sizes)], min(c for c in cuts if c > t) - t)
            k += 1
            rc, done = conv.Process(xd[:, t * B:], yd[:, t * B:], n * B)
            assert rc == T.AMF_OK and done == n * B
            t += n
        conv.Synchronize()
        conv.Terminate()
        err = np.abs(yd.cpu().numpy().astype(np.float64) - yo).max()
        print(f"K={K} phase {phase} alone: max|gpu-oracle| = {err:.3e}", flush=True)
        assert err <= 1e-5
    ctx = T.TANContext(device=0)
    objs = []
    for phase in range(K):
        c = T.TANConvolution(ctx, pass_blocks=K, pass_phase=phase)
        assert c.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
        objs.append(c)
    ys = [torch.zeros_like(xd) for _ in objs]
    assert ctx.EnableGroupTiming(True) == T.AMF_OK
    pass_bytes = C * (L // B - 2) * B * 16
    head_max = K * (C * B * 44 + (K - 1) * C * B * 16)  # K heads, up to K - 1 further partitions
    for t in range(nb):
        if t in irs:
            h = torch.from_numpy(irs[t]).cuda()
            for c in objs:
                assert c.UpdateResponseTD(h, L, None, BLOCKING) == T.AMF_OK
        assert T.GroupProcess(objs, [xd[:, t * B:]] * K, [y[:, t * B:] for y in ys]) == T.AMF_OK
        tb = ctx.ReadGroupTiming(reset=True)["tick_alg_bytes"]
        if 30 <= t < 50:
            assert pass_bytes <= tb <= pass_bytes + head_max, (t, tb)
    torch.cuda.synchronize()
    for phase, y in enumerate(ys):
        err = np.abs(y.cpu().numpy().astype(np.float64) - yo).max()
        print(f"K={K} phase {phase} in group: max|gpu-oracle| = {err:.3e}", flush=True)
        assert err <= 1e-5
    for c in objs:
        c.Terminate()
    ctx.Terminate()


def test_two_ir_slots(gpu_ctx):
    """ir_slots=2: the same output as three slots over updates that never
    fall on an armed switch; an update while a switch is armed (after the
    block that hands the previous one off and the block that arms it) returns
    AMF_INPUT_FULL and changes nothing; non-uniform and time-domain objects
    refuse two slots."""
    import torch
    L, B, C, nb = 16384, 256, 2, 90
    x = T.signals.input_bank(C, nb * B, base_seed=81)
    xd = torch.from_numpy(x).cuda()
    ups = [0, 3, 9, 10, 30, 31, 47, 60]   # 10 and 31: replace a pending update
    irs = {b: torch.from_numpy(T.signals.ir_bank(C, L, k, base_seed=82)).cuda() for k, b in enumerate(ups)}
    outs = []
    for slots in (3, 2):
        conv = T.TANConvolution(gpu_ctx, ir_slots=slots, pass_blocks=8)
        assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
        assert conv.Info()["ir_slots"] == slots
        yd = torch.zeros_like(xd)
        for b in range(nb):
            if b in irs:
                assert conv.UpdateResponseTD(irs[b], L, None, BLOCKING) == T.AMF_OK
            rc, done = conv.Process(xd[:, b * B:], yd[:, b * B:], B)
            assert rc == T.AMF_OK and done == B
        conv.Synchronize()
        outs.append(yd.cpu().numpy())
        conv.Terminate()
    np.testing.assert_array_equal(outs[1], outs[0])
    yo = _oracle_run(L, B, x, {b: irs[b].cpu().numpy() for b in irs}, nb)
    assert np.abs(outs[1].astype(np.float64) - yo).max() <= 1e-5
    # the armed window: update before block 5, blocks 5 and 6 hand it off and arm it
    conv = T.TANConvolution(gpu_ctx, ir_slots=2)
    twin = T.TANConvolution(gpu_ctx)
    for c in (conv, twin):
        assert c.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
    y2, y3 = torch.zeros_like(xd), torch.zeros_like(xd)
    for b in range(12):
        if b == 5:
            for c in (conv, twin):
                assert c.UpdateResponseTD(irs[3], L, None, BLOCKING) == T.AMF_OK
        if b == 7:
            # three slots take it (into the third), two refuse it
            assert conv.UpdateResponseTD(irs[9], L, None, BLOCKING) == T.AMF_INPUT_FULL
            assert twin.UpdateResponseTD(irs[9], L, None, BLOCKING) == T.AMF_OK
        assert conv.Process(xd[:, b * B:], y2[:, b * B:], B)[0] == T.AMF_OK
        assert twin.Process(xd[:, b * B:], y3[:, b * B:], B)[0] == T.AMF_OK
    conv.Synchronize()
    twin.Synchronize()
    assert conv.Info()["crossfades"] == 1 and twin.Info()["crossfades"] == 2
    # up to the twin's second switch the two are the same
    np.testing.assert_array_equal(y2[:, :9 * B].cpu().numpy(), y3[:, :9 * B].cpu().numpy())
    # once the crossfade block has run, the slot is free again
    assert conv.UpdateResponseTD(irs[9], L, None, BLOCKING) == T.AMF_OK
    for c in (conv, twin):
        c.Terminate()
    nu = T.TANConvolution(gpu_ctx, ir_slots=2)
    assert nu.InitGpu(T.TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM, 65536, 256, 2) == T.AMF_NOT_SUPPORTED
    td = T.TANConvolution(gpu_ctx, ir_slots=2)
    assert td.InitGpu(T.TAN_CONVOLUTION_METHOD_TIME_DOMAIN, 256, 256, 2) == T.AMF_NOT_SUPPORTED
    bad = T.TANConvolution(gpu_ctx, ir_slots=4)
    assert bad.InitGpu(UNIFORM, L, B, C) == T.AMF_INVALID_ARG
