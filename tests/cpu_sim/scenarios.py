"""Launch sequences of the GPU tests replayed on the CPU simulation build
(memory safety and barrier checks before they reach a GPU).

Run one scenario per process: `python scenarios.py <name>`; exit status 0
when it completes, 86 when an access leaves its allocation (the simulator's
trap), 87 on a barrier deadlock.  Numerics are checked against the oracle
where the scenario is small enough.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))
import simlib  # noqa: E402

lib = simlib.load()
import oracle_lib as O  # noqa: E402
import tan_amd as T  # noqa: E402
from tan_amd import batching as Bt  # noqa: E402

UNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
_done = ctypes.c_size_t(0)


def process(conv, xd, yd, t, nblocks, B):
    rc = lib.tanx_conv_process_device(conv.handle, xd.at(t * B), xd.stride, yd.at(t * B), yd.stride, nblocks * B,
                                      None, ctypes.byref(_done))
    assert rc == T.AMF_OK and _done.value == nblocks * B, (rc, _done.value)


def oracle(L, B, x, updates, nb):
    C = x.shape[0]
    orc = O.OracleConvolution(L, B, C)
    yo = np.zeros((C, nb * B), np.float32)
    for b in range(nb):
        if b in updates:
            assert orc.update_td(np.ascontiguousarray(updates[b])) == O.AMF_OK
        yo[:, b * B:(b + 1) * B] = orc.process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]))[1]
    return yo


def check(y, yo, label, tol=1e-5):
    err = float(np.abs(y.astype(np.float64) - yo).max())
    print(f"{label}: max|sim-oracle| = {err:.3e}", flush=True)
    assert err <= tol, (label, err)


def headline_calls(nb=120, C=64, L=65536, B=256, every=50):
    """Calls of up to 50 blocks on a 64-channel object with an IR update
    before each call (the last call is shorter), batched heads beside the
    multi-block passes; against the oracle on 4 channels."""
    ctx = T.TANContext(device=0)
    conv = T.TANConvolution(ctx)
    assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
    x = T.signals.input_bank(C, nb * B, base_seed=501)
    xd = simlib.Dev(C, nb * B, x)
    yd = simlib.Dev(C, nb * B, 0.0)
    ups = {}
    for t in range(0, nb, every):
        ups[t] = T.signals.ir_bank(C, L, t // every, base_seed=502)
        bank = simlib.Dev(C, L, ups[t])
        assert lib.tanx_conv_update_td_device(conv.handle, bank.ptr, bank.stride, L, None, 0) == T.AMF_OK
        process(conv, xd, yd, t, min(every, nb - t), B)
    check(yd.a[:4].copy(), oracle(L, B, x[:4], {b: h[:4] for b, h in ups.items()}, nb), "headline calls")


def headline_overrun(C=4, L=4096, B=256):
    """A planted mistake: a call asking for more blocks than the buffers
    hold.  The simulator must trap it (exit status 86)."""
    ctx = T.TANContext(device=0)
    conv = T.TANConvolution(ctx)
    assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
    xd = simlib.Dev(C, 20 * B, 0.5)
    yd = simlib.Dev(C, 20 * B, 0.0)
    process(conv, xd, yd, 0, 16, B)
    process(conv, xd, yd, 16, 8, B)  # 4 blocks past the end of the last row


def streaming_pass(nb=60, C=256, L=65536, B=256):
    """An object above the cached bound (nontemporal pass loads), calls of
    uneven size and updates at uneven offsets, one while a switch is armed."""
    ctx = T.TANContext(device=0)
    conv = T.TANConvolution(ctx)
    assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
    info = conv.Info()
    assert info["pass_blocks"] == 4 and info["cached_rows"] == 0, info
    rng = np.random.default_rng(611)
    x = rng.uniform(-1, 1, (C, nb * B)).astype(np.float32)
    xd = simlib.Dev(C, nb * B, x)
    yd = simlib.Dev(C, nb * B, 0.0)
    upd = [0, 23, 37, 38]
    banks = {b: simlib.Dev(C, L, T.signals.ir_bank(C, L, k, base_seed=612)) for k, b in enumerate(upd)}
    sizes = [1, 3, 7, 2, 13, 4, 1, 1, 9, 5]
    cuts = sorted(set(upd + [nb]))
    t = k = 0
    while t < nb:
        if t in banks:
            assert lib.tanx_conv_update_td_device(conv.handle, banks[t].ptr, L, L, None, 0) == T.AMF_OK
        n = min(sizes[k % len(sizes)], min(c for c in cuts if c > t) - t)
        k += 1
        process(conv, xd, yd, t, n, B)
        t += n
    check(yd.a[:4].copy(), oracle(L, B, x[:4], {b: banks[b].a[:4].copy() for b in upd}, nb), "streaming pass")


def c3_group(nticks=300, nstreams=64):
    """Group launches of the C3 draw (512 streams; every eighth one here,
    phases from the full set) with inputs and outputs resident, as the C3
    test drives them; two streams against the oracle."""
    specs = [(L, B, 1) for L, B in Bt.draw_streams(512, seed=7)]
    base, periods, phases, cycle = Bt.phase_plan([b for _, b, _ in specs])
    sel = list(range(0, 512, 512 // nstreams))
    ctx = T.TANContext(device=0)
    convs, irs, xs, ys = {}, {}, {}, {}
    for j in sel:
        L, B, C = specs[j]
        c = T.TANConvolution(ctx)
        assert c.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
        irs[j] = T.signals.ir_bank(C, L, 0, base_seed=900 + j)
        dev_ir = simlib.Dev(C, L, irs[j])
        assert lib.tanx_conv_update_td_device(c.handle, dev_ir.ptr, L, L, None, 1) == T.AMF_OK
        convs[j] = c
        n = len(range(phases[j], nticks, periods[j]))
        xs[j] = simlib.Dev(C, n * B, np.random.default_rng(j).uniform(-1, 1, (C, n * B)))
        ys[j] = simlib.Dev(C, n * B, 0.0)
    due = [[j for j in sel if t % periods[j] == phases[j]] for t in range(cycle)]
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    for t in range(nticks):
        d = due[t % cycle]
        blk = [(t - phases[j]) // periods[j] for j in d]
        hs = (vp * len(d))(*[convs[j].handle.value for j in d])
        ins = (vp * len(d))(*[xs[j].at(specs[j][1] * k) for j, k in zip(d, blk)])
        outs = (vp * len(d))(*[ys[j].at(specs[j][1] * k) for j, k in zip(d, blk)])
        st = (sz * len(d))(*[xs[j].stride for j in d])
        assert lib.tanx_group_process_device(len(d), hs, ins, st, outs, st, None) == T.AMF_OK
    for j in (sel[1], sel[-1]):
        L, B, C = specs[j]
        nb = xs[j].shape[1] // B
        check(ys[j].a.copy(), oracle(L, B, xs[j].a.copy(), {0: irs[j]}, nb), f"C3 stream {j} (L={L}, B={B})")


# sessions of the long serving test: (channels, IR taps, first tick, last tick + 1, update ticks)
LONG_PLAN = {
    0: (2, 131072, 0, 2200, (500, 1300)),
    1: (1, 524288, 4, 2200, (700,)),
    2: (3, 200000, 40, 1600, (90, 900)),
    3: (1, 8192, 0, 300, (10,)),
    4: (5, 300000, 900, 2200, (37,)),
}


def _serving_long(method, nticks, B=256, kp=0):
    """The long serving test's sessions arriving, updating and leaving, all
    active ones advanced in one group launch per tick (uniform ones with
    pass length kp and phase sid % kp when kp is given)."""
    ctx = T.TANContext(device=0)
    live, xs, ys, irs, done_ticks = {}, {}, {}, {}, {}
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    for t in range(nticks):
        for sid, (C, L, t0, t1, ups) in LONG_PLAN.items():
            if t == t1 and sid in live:
                live.pop(sid).Terminate()
            if t == t0:
                c = T.TANConvolution(ctx, pass_blocks=kp, pass_phase=sid % kp if kp else 0)
                assert c.InitGpu(method, L, B, C) == T.AMF_OK
                nb = min(t1, nticks) - t0
                xs[sid] = simlib.Dev(C, nb * B, np.random.default_rng(sid).uniform(-1, 1, (C, nb * B)))
                ys[sid] = simlib.Dev(C, nb * B, 0.0)
                irs[sid] = [T.signals.ir_bank(C, L, k, base_seed=300 + sid) for k in range(len(ups) + 1)]
                live[sid] = c
            if sid in live and (t - t0 == 0 or (t - t0) in ups):
                k = 0 if t == t0 else 1 + list(ups).index(t - t0)
                dev = simlib.Dev(C, L, irs[sid][k])
                assert lib.tanx_conv_update_td_device(live[sid].handle, dev.ptr, L, L, None, 0) == T.AMF_OK
                dev.free()
        act = list(live)
        hs = (vp * len(act))(*[live[s].handle.value for s in act])
        ins = (vp * len(act))(*[xs[s].at(B * (t - LONG_PLAN[s][2])) for s in act])
        outs = (vp * len(act))(*[ys[s].at(B * (t - LONG_PLAN[s][2])) for s in act])
        ist = (sz * len(act))(*[xs[s].stride for s in act])
        assert lib.tanx_group_process_device(len(act), hs, ins, ist, outs, ist, None) == T.AMF_OK
        if t % 200 == 0:
            print(f"tick {t}: {len(act)} sessions", flush=True)
    return xs, ys, irs


def serving_long_uniform(nticks=2200, kp=0):
    xs, ys, irs = _serving_long(UNIFORM, nticks, kp=kp)
    for sid in (1, 3):
        C, L, t0, t1, ups = LONG_PLAN[sid]
        nb = xs[sid].shape[1] // 256
        upd = {0: irs[sid][0]}
        upd.update({b: irs[sid][k + 1] for k, b in enumerate(ups)})
        check(ys[sid].a.copy(), oracle(L, 256, xs[sid].a.copy(), upd, nb), f"serving session {sid} (L={L})")


def serving_long_nonuniform(nticks=2200):
    _serving_long(T.TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM, nticks)


def noalloc_group():
    """The allocation-free group entry point beside the allocating one on twin
    objects (the sequence of tests/test_zz_no_alloc.py): bit-identical
    outputs, calls whose object count drops and rises."""
    specs = [(65536, 256, 4, UNIFORM), (5000, 128, 2, UNIFORM), (131072, 512, 1, UNIFORM),
             (32768, 256, 2, T.TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM), (4096, 64, 3, UNIFORM)]
    calls = [[0, 1, 2, 3, 4], [1, 3, 4]] * 24
    ctx = T.TANContext(device=0)
    irs = [simlib.Dev(C, L, T.signals.ir_bank(C, L, 0, base_seed=77 + i)) for i, (L, B, C, m) in enumerate(specs)]
    sets = []
    for _ in range(2):
        objs = []
        for (L, B, C, m), h in zip(specs, irs):
            c = T.TANConvolution(ctx)
            assert c.InitGpu(m, L, B, C) == T.AMF_OK
            assert lib.tanx_conv_update_td_device(c.handle, h.ptr, L, L, None, 1) == T.AMF_OK
            objs.append(c)
        sets.append(objs)
    nblk = [sum(1 for c in calls if i in c) for i in range(len(specs))]
    xs = [simlib.Dev(C, n * B, np.random.default_rng(i).uniform(-1, 1, (C, n * B)))
          for i, ((L, B, C, m), n) in enumerate(zip(specs, nblk))]
    ys = [[simlib.Dev(C, n * B, 0.0) for (L, B, C, m), n in zip(specs, nblk)] for _ in range(2)]
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    blk = [0] * len(specs)
    for idx in calls:
        for fn, objs, y in ((lib.tanx_group_process_device, sets[0], ys[0]),
                            (lib.tanx_group_process_device_noalloc, sets[1], ys[1])):
            n = len(idx)
            hs = (vp * n)(*[objs[i].handle.value for i in idx])
            ins = (vp * n)(*[xs[i].at(specs[i][1] * blk[i]) for i in idx])
            outs = (vp * n)(*[y[i].at(specs[i][1] * blk[i]) for i in idx])
            st = (sz * n)(*[xs[i].stride for i in idx])
            assert fn(n, hs, ins, st, outs, st, None) == T.AMF_OK
        for i in idx:
            blk[i] += 1
    for i in range(len(specs)):
        assert np.array_equal(ys[0][i].a, ys[1][i].a), i
        assert float(np.abs(ys[1][i].a).max()) > 0.01, i
    many = (vp * 1025)()
    szs = (sz * 1025)()
    assert lib.tanx_group_process_device_noalloc(1025, many, many, szs, many, szs, None) == T.AMF_INVALID_ARG
    print("noalloc group: outputs bit-identical", flush=True)


def queue_group(ncalls=40):
    """Work-queue group launches (tanx_ctx_set_group_queue) beside the plain
    ones on twin objects of two contexts: multi-block passes of several
    lengths, a non-uniform object, IR updates with their crossfades, calls
    whose object count drops and rises; bit-identical outputs."""
    NU = T.TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM
    specs = [(65536, 256, 4, UNIFORM, 4), (5000, 128, 2, UNIFORM, 0), (131072, 512, 1, UNIFORM, 8),
             (32768, 256, 2, NU, 0), (16384, 64, 3, UNIFORM, 12), (65536, 256, 3, UNIFORM, 12)]
    calls = [[0, 1, 2, 3, 4, 5], [1, 3, 4, 5]] * (ncalls // 2)
    irs = [[simlib.Dev(C, L, T.signals.ir_bank(C, L, u, base_seed=91 + i)) for u in range(2)]
           for i, (L, B, C, m, kp) in enumerate(specs)]
    sets = []
    for q in range(2):
        ctx = T.TANContext(device=0)
        assert ctx.SetGroupQueue(q == 1) == T.AMF_OK
        objs = []
        for i, (L, B, C, m, kp) in enumerate(specs):
            kw = {"pass_blocks": kp, "pass_phase": i % kp} if kp else {}
            c = T.TANConvolution(ctx, **kw)
            assert c.InitGpu(m, L, B, C) == T.AMF_OK
            assert lib.tanx_conv_update_td_device(c.handle, irs[i][0].ptr, L, L, None, 1) == T.AMF_OK
            objs.append(c)
        sets.append(objs)
    nblk = [sum(1 for c in calls if i in c) for i in range(len(specs))]
    xs = [simlib.Dev(C, n * B, np.random.default_rng(i).uniform(-1, 1, (C, n * B)))
          for i, ((L, B, C, m, kp), n) in enumerate(zip(specs, nblk))]
    ys = [[simlib.Dev(C, n * B, 0.0) for (L, B, C, m, kp), n in zip(specs, nblk)] for _ in range(2)]
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    blk = [0] * len(specs)
    for t, idx in enumerate(calls):
        for objs, y in zip(sets, ys):
            if t % 9 == 4:
                # a fresh IR for two objects (crossfades in the next calls)
                for i in (0, 5):
                    L = specs[i][0]
                    rc = lib.tanx_conv_update_td_device(objs[i].handle, irs[i][(t // 9) % 2 ^ 1].ptr, L, L, None, 0)
                    assert rc == T.AMF_OK, rc
            n = len(idx)
            hs = (vp * n)(*[objs[i].handle.value for i in idx])
            ins = (vp * n)(*[xs[i].at(specs[i][1] * blk[i]) for i in idx])
            outs = (vp * n)(*[y[i].at(specs[i][1] * blk[i]) for i in idx])
            st = (sz * n)(*[xs[i].stride for i in idx])
            assert lib.tanx_group_process_device(n, hs, ins, st, outs, st, None) == T.AMF_OK
        for i in idx:
            blk[i] += 1
    for i in range(len(specs)):
        assert np.array_equal(ys[0][i].a, ys[1][i].a), i
        assert float(np.abs(ys[1][i].a).max()) > 0.01, i
    print("queue group: outputs bit-identical", flush=True)


def batched_calls(kp=8, B=256, precision=0):
    """Calls of several blocks with K-block passes (heads batched with the
    next blocks' pass, head-only batches, catch-up passes after updates)
    and dense updates between calls; against the oracle (fp32 spectra) or
    against one block per call."""
    calls = [1, 3, 7, 2, 13, 5, 1, 1, 9, 4, 6, 11, 2, 8]
    L, C = 32 * B, 3
    nb = sum(calls)
    starts = np.cumsum([0] + calls[:-1])
    ups = {int(starts[i]): T.signals.ir_bank(C, L, update=k, base_seed=601)
           for k, i in enumerate((0, 2, 3, 5, 6, 7, 10, 12))}
    x = T.signals.input_bank(C, nb * B, base_seed=700)
    ctx = T.TANContext(device=0)
    ys = []
    for sizes in (calls, [1] * nb):
        conv = T.TANConvolution(ctx, ir_precision=precision, pass_blocks=kp)
        assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
        xd = simlib.Dev(C, nb * B, x)
        yd = simlib.Dev(C, nb * B, 0.0)
        t = 0
        for n in sizes:
            if t in ups:
                dev = simlib.Dev(C, L, ups[t])
                assert lib.tanx_conv_update_td_device(conv.handle, dev.ptr, L, L, None, 0) == T.AMF_OK
            process(conv, xd, yd, t, n, B)
            t += n
        ys.append(yd.a.copy())
    err = float(np.abs(ys[0].astype(np.float64) - ys[1]).max())
    print(f"K={kp} B={B}: calls vs one block per call {err:.3e}", flush=True)
    assert err <= 1e-6
    if precision == 0:
        check(ys[0], oracle(L, B, x, ups, nb), f"K={kp} B={B} calls")


def fft_large(log2n=13, C=2):
    """TANFFT's four-step sizes, every direction, host buffers and device
    buffers in place, against numpy's FP64 FFT."""
    ctx = T.TANContext(device=0)
    fft = T.TANFFT(ctx)
    assert fft.Init() == T.AMF_OK
    n, m = 1 << log2n, 1 << (log2n - 1)
    rng = np.random.default_rng(log2n)
    z = (rng.standard_normal((C, n)) + 1j * rng.standard_normal((C, n))).astype(np.complex64)
    buf = np.zeros((C, 2 * n), np.float32)
    buf[:, 0::2], buf[:, 1::2] = z.real, z.imag
    out = np.zeros_like(buf)
    assert fft.Transform(0, log2n, C, buf, out) == T.AMF_OK
    ref = np.fft.fft(z.astype(np.complex128), axis=1)
    err = np.abs(out[:, 0::2] + 1j * out[:, 1::2] - ref).max() / np.abs(ref).max()
    print(f"complex 2^{log2n}: relative error {err:.2e}", flush=True)
    assert err <= 3e-7 * log2n
    back = np.zeros_like(buf)
    assert fft.Transform(1, log2n, C, out, back) == T.AMF_OK
    assert np.abs(back - buf).max() <= 1.5e-6 * log2n
    # real directions: interleaved and planar, and back
    x = rng.standard_normal((C, n)).astype(np.float32)
    rb = np.zeros((C, 2 * n), np.float32)
    rb[:, :n] = x
    rref = np.fft.rfft(x.astype(np.float64), axis=1)
    for fwd, inv in ((2, 3), (4, 5)):
        o = np.full_like(rb, 9.0)
        assert fft.Transform(fwd, log2n, C, rb, o) == T.AMF_OK
        got = (o[:, 0:n + 2:2] + 1j * o[:, 1:n + 2:2]) if fwd == 2 else (o[:, :m + 1] + 1j * o[:, m + 8:2 * m + 9])
        err = np.abs(got - rref).max() / np.abs(rref).max()
        print(f"real 2^{log2n} mode {fwd}: relative error {err:.2e}", flush=True)
        assert err <= 3e-7 * log2n
        bk = np.zeros_like(rb)
        assert fft.Transform(inv, log2n, C, o, bk) == T.AMF_OK
        assert np.abs(bk[:, :n] - x).max() <= 1e-6 * log2n
    # device buffer, in place, through the C ABI
    dev = simlib.Dev(C, 2 * n, buf)
    assert lib.tanx_fft_transform_device(ctx.handle, 0, log2n, C, dev.ptr, dev.ptr, dev.stride) == T.AMF_OK
    assert np.array_equal(dev.a, out)
    print("four-step transforms complete", flush=True)


GROUP_UPDATE_SPECS = [  # (ir_len, block, channels, method, ir precision)
    (65536, 256, 3, UNIFORM, 0),
    (5000, 128, 2, UNIFORM, 0),
    (131072, 512, 1, UNIFORM, 0),
    (32768, 256, 2, T.TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM, 0),
    (20000, 256, 2, UNIFORM, 1),
    (4096, 64, 3, UNIFORM, 2),
]


def group_updates(nb=40, every=7):
    """IR updates of several objects in one grouped call beside the same
    updates one object at a time on twin objects (block sizes 64-512,
    non-uniform frames, FP16 and BF16 spectra, a channel held by
    STOP_INPUT): the outputs are the same bits."""
    ctx = T.TANContext(device=0)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    twins = []
    for L, Bk, C, m, prec in GROUP_UPDATE_SPECS:
        pair = []
        for _ in range(2):
            c = T.TANConvolution(ctx, ir_precision=prec)
            assert c.InitGpu(m, L, Bk, C) == T.AMF_OK
            pair.append(c)
        x = simlib.Dev(C, nb * Bk, np.random.default_rng(L).uniform(-1, 1, (C, nb * Bk)))
        ys = [simlib.Dev(C, nb * Bk, 0.0) for _ in range(2)]
        twins.append((pair, x, ys, L, Bk, C))
    irs = [[simlib.Dev(C, L, T.signals.ir_bank(C, L, k, base_seed=70 + i)) for k in range(3)]
           for i, (_, _, _, L, _, C) in enumerate(twins)]
    flags = [np.zeros(C, np.uint32) for (_, _, _, _, _, C) in twins]
    flags[0][1] = T.TAN_CONVOLUTION_CHANNEL_FLAG_STOP_INPUT
    for t in range(nb):
        if t % every == 0:
            k = (t // every) % 3
            fl = [f if t else None for f in flags]  # the first update sets every channel
            for i, (pair, x, ys, L, Bk, C) in enumerate(twins):
                fp = fl[i].ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)) if fl[i] is not None else None
                assert lib.tanx_conv_update_td_device(pair[0].handle, irs[i][k].ptr, L, L, fp, 0) == T.AMF_OK
            n = len(twins)
            hs = (vp * n)(*[p[1].handle.value for p, *_ in twins])
            ps = (vp * n)(*[irs[i][k].ptr for i in range(n)])
            st = (sz * n)(*[tw[3] for tw in twins])
            fa = (vp * n)(*[f.ctypes.data if f is not None else None for f in fl])
            assert lib.tanx_group_update_td_device(n, hs, ps, st, st, ctypes.cast(fa, vp), 0) == T.AMF_OK
        for pair, x, ys, L, Bk, C in twins:
            for j in range(2):
                process(pair[j], x, ys[j], t, 1, Bk)
    for i, (pair, x, ys, L, Bk, C) in enumerate(twins):
        assert np.array_equal(ys[0].a, ys[1].a), i
        assert float(np.abs(ys[0].a).max()) > 0.01, i
    bad = (vp * 2)(twins[0][0][1].handle.value, twins[0][0][1].handle.value)
    st = (sz * 2)(65536, 65536)
    assert lib.tanx_group_update_td_device(2, bad, (vp * 2)(irs[0][0].ptr, irs[0][0].ptr), st, st, None,
                                           0) == T.AMF_INVALID_ARG
    print("grouped updates: outputs bit-identical to per-object updates", flush=True)


def pass_phases(C=2, L=16384, B=128, nb=80, kp=4):
    """Objects with pass length kp and phases 0..kp-1: alone with calls of
    uneven size and dense updates, and together in group launches (one block
    each per call); every channel against the oracle."""
    ctx = T.TANContext(device=0)
    x = T.signals.input_bank(C, nb * B, base_seed=41)
    ups = {0: 0, 9: 1, 10: 2, 23: 3, 24: 4, 25: 5, 50: 6}
    irs = {b: T.signals.ir_bank(C, L, k, base_seed=42) for b, k in ups.items()}
    yo = oracle(L, B, x, irs, nb)
    sizes = [1, 2, 5, 3, 7, 1, 4]
    for phase in range(kp):
        conv = T.TANConvolution(ctx, pass_blocks=kp, pass_phase=phase)
        assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
        xd = simlib.Dev(C, nb * B, x)
        yd = simlib.Dev(C, nb * B, 0.0)
        t = k = 0
        cuts = sorted(set(list(irs) + [nb]))
        while t < nb:
            if t in irs:
                dev = simlib.Dev(C, L, irs[t])
                assert lib.tanx_conv_update_td_device(conv.handle, dev.ptr, L, L, None, 0) == T.AMF_OK
            n = min(sizes[k % len(sizes)], min(c for c in cuts if c > t) - t)
            k += 1
            process(conv, xd, yd, t, n, B)
            t += n
        check(yd.a.copy(), yo, f"phase {phase} alone")
    # kp phased objects in one group launch per block
    objs = []
    for phase in range(kp):
        c = T.TANConvolution(ctx, pass_blocks=kp, pass_phase=phase)
        assert c.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
        objs.append(c)
    xd = simlib.Dev(C, nb * B, x)
    ys = [simlib.Dev(C, nb * B, 0.0) for _ in objs]
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    assert ctx.EnableGroupTiming(True) == T.AMF_OK
    P = L // B
    pass_bytes = C * (P - 2) * B * 16          # FDL + IR rows of partitions 2..P-1
    head_max = kp * (C * B * 44 + (kp - 1) * C * B * 16)  # kp heads with up to kp - 1 further partitions
    for t in range(nb):
        if t in irs:
            dev = simlib.Dev(C, L, irs[t])
            for c in objs:
                assert lib.tanx_conv_update_td_device(c.handle, dev.ptr, L, L, None, 0) == T.AMF_OK
        n = len(objs)
        hs = (vp * n)(*[c.handle.value for c in objs])
        ins = (vp * n)(*[xd.at(t * B)] * n)
        outs = (vp * n)(*[y.at(t * B) for y in ys])
        st = (sz * n)(*[xd.stride] * n)
        assert lib.tanx_group_process_device(n, hs, ins, st, outs, st, None) == T.AMF_OK
        tb = ctx.ReadGroupTiming(reset=True)["tick_alg_bytes"]
        if 30 <= t < 50:  # no switch in flight: exactly one object's pass per tick
            assert pass_bytes <= tb <= pass_bytes + head_max, (t, tb)
    for phase, y in enumerate(ys):
        check(y.a.copy(), yo, f"phase {phase} in group")


def rt_rig(groups=8, C=16, L=65536, B=256, nticks=130, kp=4):
    """The benchmark's real-time rig at a small size: groups with pass length
    kp and phases g % kp, each group's IR update at its own offset in the 50-block period
    (transform on the process stream before that tick's launch), one group
    launch per tick with group timing on; two groups against the oracle."""
    ctx = T.TANContext(device=0)
    assert ctx.EnableGroupTiming(True) == T.AMF_OK
    banks = [simlib.Dev(C, L, T.signals.ir_bank(C, L, i, base_seed=5000)) for i in range(4)]
    objs = []
    for g in range(groups):
        c = T.TANConvolution(ctx, pass_blocks=kp, pass_phase=g % kp)
        assert c.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
        assert lib.tanx_conv_update_td_device(c.handle, banks[g % 4].ptr, L, L, None, 0) == T.AMF_OK
        objs.append(c)
    x = np.random.default_rng(11).uniform(-1, 1, (groups * C, nticks * B)).astype(np.float32)
    xd = simlib.Dev(groups * C, nticks * B, x)
    yd = simlib.Dev(groups * C, nticks * B, 0.0)
    ups = [dict() for _ in range(groups)]
    for g in range(groups):
        ups[g][0] = banks[g % 4].a.copy()
    nupd = [0] * groups
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    for t in range(nticks):
        for g in range(groups):
            if (t + (g * 50) // groups) % 50 == 0:
                nupd[g] += 1
                bank = banks[(nupd[g] + g) % 4]
                assert lib.tanx_conv_update_td_device(objs[g].handle, bank.ptr, L, L, None, 0) == T.AMF_OK
                ups[g][t] = bank.a.copy()
        hs = (vp * groups)(*[o.handle.value for o in objs])
        ins = (vp * groups)(*[xd.at(t * B) + 4 * g * C * xd.stride for g in range(groups)])
        outs = (vp * groups)(*[yd.at(t * B) + 4 * g * C * yd.stride for g in range(groups)])
        st = (sz * groups)(*[xd.stride] * groups)
        assert lib.tanx_group_process_device(groups, hs, ins, st, outs, st, None) == T.AMF_OK
    gt = ctx.ReadGroupTiming(reset=True)
    assert gt["tick_launches"] == nticks and gt["objects"] == nticks * groups, gt
    for g in (1, groups - 1):
        sl = slice(g * C, g * C + 2)
        check(yd.a[sl].copy(), oracle(L, B, x[sl], {b: h[:2] for b, h in ups[g].items()}, nticks), f"rt group {g}")


if __name__ == "__main__":
    # python scenarios.py NAME [key=int ...]
    globals()[sys.argv[1]](**{k: int(v) for k, v in (a.split("=") for a in sys.argv[2:])})
    print("scenario complete", flush=True)
