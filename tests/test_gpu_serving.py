"""Continuous batching: sessions that arrive, update their IR and leave while
others run give the same output in the mix as processed alone, bit for bit,
and match the oracle; HBM-budget admission holds a session back until
another one leaves.  Long sessions (131,072-524,288 taps, run past their
last partition) are checked with the uniform method against the oracle and
with the non-uniform method against the FP64 model of the time-varying
convolution at its frame-aligned switch blocks."""
import numpy as np
import pytest

import oracle_lib as O
import tan_amd as T
from helpers import crossfade_model, snr_db
from tan_amd import serving as S
from test_gpu_nonuniform import nu_schedule

pytestmark = pytest.mark.gpu
UNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
NONUNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM
B = 256

# sid: (channels, ir_len, arrive tick, leave tick, ticks of IR updates)
PLAN = {
    0: (2, 5000, 0, 60, (8, 30)),
    1: (2, 20000, 3, 40, ()),
    2: (3, 3000, 0, 60, (15,)),
    3: (4, 12000, 10, 60, (20, 21, 44)),
    4: (16, 2048, 25, 60, (33,)),
}

# long sessions of the served workload's IR range; 1 and 4 (P = 2,048 at
# 524,288 taps) run past their last partition
LONG_PLAN = {
    0: (2, 131072, 0, 2200, (500, 1300)),
    1: (1, 524288, 4, 2200, (700,)),
    2: (3, 200000, 40, 1600, (90, 900)),
    3: (1, 8192, 0, 300, (10,)),
    4: (5, 300000, 900, 2200, (37,)),
}


def _ir(sid, k, C, L):
    return T.signals.ir_bank(C, L, k, base_seed=300 + sid)


def _x(sid, tick, C):
    rng = np.random.default_rng([sid, tick])
    return rng.uniform(-1.0, 1.0, (C, B)).astype(np.float32)


def _run_mix(ctx, budget, plan=PLAN, nticks=60, method=UNIFORM, **server_kw):
    """Drive a server through the plan; returns {sid: (admit tick, outputs,
    the object's options)}."""
    import torch
    srv = S.SessionServer(ctx, block=B, hbm_budget=budget, method=method, **server_kw)
    outs = {sid: [] for sid in plan}
    admitted_at, opts = {}, {}
    upd_count = {sid: 0 for sid in plan}
    for t in range(nticks):
        for sid, (C, L, t0, t1, ups) in plan.items():
            if t == t1 and (sid in srv.active or sid in srv.pending):
                srv.release(sid)
            if t == t0:
                assert srv.submit(sid, C, L, torch.from_numpy(_ir(sid, 0, C, L)).cuda())
        for s in srv.admit():
            admitted_at[s.sid] = t
            opts[s.sid] = s.conv_options()
        for sid, (C, L, t0, t1, ups) in plan.items():
            if sid in srv.active and (t - admitted_at[sid]) in ups:
                upd_count[sid] += 1
                srv.update(sid, torch.from_numpy(_ir(sid, upd_count[sid], C, L)).cuda())
        for sid, s in srv.active.items():
            s.inp.copy_(torch.from_numpy(_x(sid, t - admitted_at[sid], s.channels)))
        rec = S.SessionServer.finish(srv.tick())
        assert rec.sessions == len(srv.active)
        for sid, s in srv.active.items():
            outs[sid].append(s.out.cpu().numpy().copy())
    srv.shutdown()
    assert not any(srv.pass_load), "released sessions return their pass bytes"
    return {sid: (admitted_at[sid], np.concatenate(o, axis=1), opts[sid]) for sid, o in outs.items() if o}


def _alone(ctx, sid, nblocks, plan=PLAN, method=UNIFORM, xfades=None, info=None, opts=None):
    """The same session on its own object (created with the server's options
    for it), one Process per block; appends the crossfade blocks to xfades
    and the object's Info() to info."""
    import torch
    C, L, t0, t1, ups = plan[sid]
    conv = T.TANConvolution(ctx, **(opts or {}))
    assert conv.InitGpu(method, L, B, C) == T.AMF_OK
    if info is not None:
        info.update(conv.Info())
    y = np.zeros((C, nblocks * B), np.float32)
    k = 0
    assert conv.UpdateResponseTD(torch.from_numpy(_ir(sid, 0, C, L)).cuda(), L) == T.AMF_OK
    yo = torch.zeros(C, B, device="cuda")
    for b in range(nblocks):
        if b in ups:
            k += 1
            assert conv.UpdateResponseTD(torch.from_numpy(_ir(sid, k, C, L)).cuda(), L) == T.AMF_OK
        xb = torch.from_numpy(_x(sid, b, C)).cuda()
        rc, done = conv.Process(xb, yo, B)
        assert rc == T.AMF_OK and done == B
        conv.Synchronize()
        y[:, b * B:(b + 1) * B] = yo.cpu().numpy()
        if xfades is not None and conv.LastWasCrossfade():
            xfades.append(b)
    conv.Terminate()
    return y


def _oracle(sid, nblocks, plan=PLAN):
    C, L, t0, t1, ups = plan[sid]
    orc = O.OracleConvolution(L, B, C)
    orc.update_td(_ir(sid, 0, C, L))
    y = np.zeros((C, nblocks * B), np.float32)
    k = 0
    for b in range(nblocks):
        if b in ups:
            k += 1
            orc.update_td(_ir(sid, k, C, L))
        _, o, _ = orc.process(_x(sid, b, C))
        y[:, b * B:(b + 1) * B] = o
    return y


def test_session_alone_equals_mix_and_oracle(gpu_ctx):
    mix = _run_mix(gpu_ctx, budget=64 << 30)
    for sid, (at, y, opts) in mix.items():
        C, L, t0, t1, ups = PLAN[sid]
        assert at == t0
        nb = y.shape[1] // B
        assert nb == t1 - t0
        ya = _alone(gpu_ctx, sid, nb, opts=opts)
        np.testing.assert_array_equal(y, ya, err_msg=f"session {sid}")
        yo = _oracle(sid, nb)
        assert np.abs(y - yo).max() <= 1e-5, f"session {sid}"
        live = [c for c in range(C) if np.any(yo[c])]
        assert live and min(snr_db(yo[c], y[c]) for c in live) >= 120.0


def test_budget_holds_session_until_release(gpu_ctx):
    probe = S.SessionServer(gpu_ctx, block=B)  # the admission charge of the servers below
    need = {sid: probe.session_bytes(C, L) for sid, (C, L, *_r) in PLAN.items()}
    # 3 (arrives at 10) does not fit beside 0, 1 and 2; 4 (arrives at 25,
    # smaller) does and goes first; 3 fits once 1 leaves at tick 40
    assert need[1] > need[4] and need[3] >= need[1]
    budget = need[0] + need[2] + need[3] + need[4]
    mix = _run_mix(gpu_ctx, budget=budget)
    at3, y3, o3 = mix[3]
    assert at3 == PLAN[1][3], "session 3 is admitted in the tick session 1 leaves"
    nb = y3.shape[1] // B
    np.testing.assert_array_equal(y3, _alone(gpu_ctx, 3, nb, opts=o3))
    for sid in (0, 1, 2, 4):
        assert mix[sid][0] == PLAN[sid][2]


def test_admission_rejects_oversized(gpu_ctx):
    import torch
    srv = S.SessionServer(gpu_ctx, block=B, hbm_budget=1 << 20)
    assert not srv.submit(0, 16, 524288, torch.zeros(16, 524288, device="cuda"))
    assert srv.admission.rejected == 1 and not srv.pending
    srv.shutdown()


def test_long_sessions_uniform_against_oracle(gpu_ctx):
    """Sessions of 131,072-524,288 taps (one arriving mid-run) over 2,200
    ticks: the mix equals each session alone, and the oracle."""
    mix = _run_mix(gpu_ctx, budget=64 << 30, plan=LONG_PLAN, nticks=2200, pass_blocks=8)
    for sid, (at, y, opts) in sorted(mix.items()):
        C, L, t0, t1, ups = LONG_PLAN[sid]
        nb = y.shape[1] // B
        assert at == t0 and nb == t1 - t0
        print(f"long uniform session {sid}: C={C} L={L} blocks={nb} {opts}", flush=True)
        np.testing.assert_array_equal(y, _alone(gpu_ctx, sid, nb, plan=LONG_PLAN, opts=opts),
                                      err_msg=f"session {sid}")
        yo = _oracle(sid, nb, plan=LONG_PLAN)
        err = float(np.abs(y.astype(np.float64) - yo).max())
        snr = min(snr_db(yo[c], y[c]) for c in range(C))
        print(f"  max|gpu-oracle|={err:.2e} snr={snr:.1f} dB", flush=True)
        assert err <= 1e-5 and snr >= 120.0, (sid, err, snr)


def test_long_sessions_nonuniform_against_fp64(gpu_ctx):
    """The same long sessions served with the non-uniform method: the mix
    equals each session alone, its crossfades fall on the frame-aligned
    schedule, and the output matches the FP64 time-varying convolution."""
    mix = _run_mix(gpu_ctx, budget=64 << 30, plan=LONG_PLAN, nticks=2200, method=NONUNIFORM)
    for sid, (at, y, opts) in sorted(mix.items()):
        C, L, t0, t1, ups = LONG_PLAN[sid]
        nb = y.shape[1] // B
        assert at == t0 and nb == t1 - t0
        xf, info = [], {}
        ya = _alone(gpu_ctx, sid, nb, plan=LONG_PLAN, method=NONUNIFORM, xfades=xf, info=info, opts=opts)
        np.testing.assert_array_equal(y, ya, err_msg=f"session {sid}")
        M = info["nu_multiple"]
        sched = nu_schedule([0] + list(ups), M, 0, nb)
        print(f"long non-uniform session {sid}: C={C} L={L} M={M} blocks={nb} crossfades {xf}", flush=True)
        assert xf == [b for b, _ in sched]
        x = np.concatenate([_x(sid, b, C) for b in range(nb)], axis=1)
        irs = [_ir(sid, k, C, L) for k in range(len(ups) + 1)]
        for c in range(C):
            ref = crossfade_model(x[c], [irs[k][c] for _, k in sched], [b for b, _ in sched], B)
            err = float(np.abs(y[c] - ref).max())
            snr = snr_db(ref, y[c])
            print(f"  ch{c} vs FP64: max abs {err:.2e}, SNR {snr:.1f} dB", flush=True)
            assert err <= 1e-5 and snr >= 120.0, (sid, c, err, snr)
