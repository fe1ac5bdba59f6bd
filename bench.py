#!/usr/bin/env python3
"""Benchmark of the time-varying multichannel convolution path on one MI355X.

Workload (the headline configuration): 64 channels, FP32, 48 kHz, block 256,
IR 65,536 taps of decaying noise (unit L2), white-noise input in [-1, 1],
2,000 blocks, a fresh IR for every channel every 50 blocks (UpdateResponseTD)
with a one-block linear crossfade.  Inputs, outputs and IR banks are resident
in HBM before the timed region; the timed region is exactly --steps blocks.

value = real-time channels sustained = channels * audio seconds / wall seconds.

Also reported (one JSON line):
  roofline      -- the dominant kernel (tail multiply-accumulate over the
                   FDL), algorithmic bytes / average launch time from HIP
                   events recorded on the process stream, against 8 TB/s;
  latency       -- p50/p99/max block latency of the real-time loop
                   (host buffers in, host buffers out, one block in flight)
                   and deadline misses against 5.333 ms;
  sweep         -- the largest channel count whose every block (one in
                   flight, device-resident) finishes within the deadline;
  cpu_baseline  -- the CPU restatement of the reference's CPU path
                   (uniform partitioned FFT convolution, planar FMA CMAC,
                   OpenMP threads = cores/4 as the reference sets them) on a
                   bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FS = 48000
B, L, C = 256, 65536, 64
UPDATE_EVERY = 50
DEADLINE_MS = 1000.0 * B / FS
HBM_PEAK_GBPS = 8000.0
METRIC = "real-time 48kHz channels @64k-tap IR, 256 blk; p99 block latency vs plain FP32"
WORKLOAD = ("plain GPU method: 64 ch FP32, block 256, IR 65536 taps (decaying noise), "
            "2000 blocks, fresh IR every 50 blocks, 1-block linear crossfade")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def ir_bank_device(torch, channels, taps, seed):
    """Decaying-noise IRs (exp(-t/0.25 s) at 48 kHz), unit L2, on the GPU."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    env = torch.exp(-torch.arange(taps, device="cuda", dtype=torch.float32) / (0.25 * FS))
    h = torch.randn(channels, taps, generator=g, device="cuda") * env
    h /= h.norm(dim=1, keepdim=True)
    return h.contiguous()


def noise_device(torch, channels, n, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return (torch.rand(channels, n, generator=g, device="cuda") * 2.0 - 1.0).contiguous()


def run_blocks(T, conv, x, y, banks, t0, t1):
    """Blocks [t0, t1) through the device-memory Process, with a new IR bank
    staged every UPDATE_EVERY blocks."""
    t = t0
    while t < t1:
        if t % UPDATE_EVERY == 0:
            rc = conv.UpdateResponseTD(banks[t // UPDATE_EVERY], L, None, 0)
            assert rc == T.AMF_OK, rc
        nb = min(UPDATE_EVERY - t % UPDATE_EVERY, t1 - t)
        rc, done = conv.Process(x[:, t * B:], y[:, t * B:], nb * B)
        assert rc == T.AMF_OK and done == nb * B, (rc, done)
        t += nb


UNIFORM_METHOD = 3  # TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
TAIL_CHUNKS = 0  # --tail-chunks: 0 lets the autotuner choose
NONUNIFORM_METHOD = 2  # TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM


def bench_throughput(T, torch, ctx, steps, warmup, timing, method=UNIFORM_METHOD, stream=None):
    """stream: the torch stream ctx was created on (its process stream); the
    inputs are written and Process is called on it."""
    if stream is not None:
        with torch.cuda.stream(stream):
            return bench_throughput(T, torch, ctx, steps, warmup, timing, method)
    conv = T.TANConvolution(ctx)
    assert conv.InitGpu(method, L, B, C) == T.AMF_OK
    if TAIL_CHUNKS:
        # pinned (a profiler slows the autotuner's dry runs and shifts its choice)
        assert conv.SetTailChunks(TAIL_CHUNKS) == T.AMF_OK
        log(f"  tail chunks pinned to {TAIL_CHUNKS}")
    else:
        rc, chunks, us = conv.Autotune()
        assert rc == T.AMF_OK, rc
        log(f"  autotune: {chunks} tail chunks, {us:.2f} us per block")
    total = warmup + steps
    nbank = total // UPDATE_EVERY + 1
    banks = [ir_bank_device(torch, C, L, 1000 + k) for k in range(nbank)]
    x = noise_device(torch, C, total * B, 7)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    run_blocks(T, conv, x, y, banks, 0, warmup)
    conv.Synchronize()
    if timing:
        conv.EnableTiming(True)
        conv.ReadTiming(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run_blocks(T, conv, x, y, banks, warmup, total)
    t_enq = time.perf_counter() - t0
    conv.Synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    log(f"  host enqueue {t_enq * 1e3:.1f} ms of {dt * 1e3:.1f} ms")
    kt = conv.ReadTiming() if timing else None
    info = conv.Info()
    ok = bool(torch.isfinite(y[:, warmup * B:]).all().item())
    conv.Terminate()
    del banks, x, y
    torch.cuda.empty_cache()
    return dt, kt, info, ok


def bench_latency(T, ctx, nblocks):
    """Real-time loop: host buffers, one block per Process call."""
    conv = T.TANConvolution(ctx)
    assert conv.InitGpu(T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED, L, B, C) == T.AMF_OK
    rng = np.random.default_rng(3)
    env = np.exp(-np.arange(L) / (0.25 * FS))
    lat = []
    xin = rng.uniform(-1, 1, (C, B)).astype(np.float32)
    out = np.zeros((C, B), np.float32)
    for b in range(nblocks):
        if b % UPDATE_EVERY == 0:
            h = (rng.standard_normal((C, L)) * env)
            h = (h / np.sqrt((h * h).sum(axis=1, keepdims=True))).astype(np.float32)
            assert conv.UpdateResponseTD(h, L, None, 0) == T.AMF_OK
        t0 = time.perf_counter()
        rc, done = conv.Process(xin, out, B)
        lat.append((time.perf_counter() - t0) * 1e3)
        assert rc == T.AMF_OK and done == B
    conv.Terminate()
    lat = np.array(lat[8:])
    worst = np.argsort(lat)[::-1][:3]
    return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "max_ms": float(lat.max()), "blocks": int(lat.size),
            "deadline_misses": int((lat > DEADLINE_MS).sum()),
            "worst": [[int(i) + 8, float(lat[i])] for i in worst]}


def bench_sweep(T, torch, ctx, counts, ticks=96, method=UNIFORM_METHOD):
    """Per-block latency with one block in flight, device-resident data."""
    from tan_amd import serving as S
    rows = []
    best = 0
    for ch in counts:
        free, _ = torch.cuda.mem_get_info()
        # the object, its IR staging and 2 x 64 blocks of input and output
        need = (S.estimate_bytes(ctx._lib, L, B, ch, method=method) + ch * L * 4 + 2 * ch * 64 * B * 4) * 1.05
        if need > free:
            log(f"sweep: skip {ch} channels (needs {need / 2**30:.1f} GiB)")
            break
        conv = T.TANConvolution(ctx)
        assert conv.InitGpu(method, L, B, ch) == T.AMF_OK
        h = ir_bank_device(torch, ch, L, 77)
        torch.cuda.synchronize()
        assert conv.UpdateResponseTD(h, L, None, T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY) == 0
        del h
        x = noise_device(torch, ch, 64 * B, 5)
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for t in range(16):
            conv.Process(x[:, (t % 64) * B:], y[:, (t % 64) * B:], B)
        conv.Synchronize()
        conv.EnableTiming(True)
        conv.ReadTiming(reset=True)
        lat = []
        for t in range(ticks):
            t0 = time.perf_counter()
            conv.Process(x[:, (t % 64) * B:], y[:, (t % 64) * B:], B)
            conv.Synchronize()
            lat.append((time.perf_counter() - t0) * 1e3)
        kt = conv.ReadTiming()
        conv.Terminate()
        del x, y
        torch.cuda.empty_cache()
        lat = np.array(lat)
        p99 = float(np.percentile(lat, 99))
        # every launch of the device path (single-block ticks, batched heads, fused head+pass), per block
        ms_all = kt["head_ms"] + kt["tail_ms"]
        tick_us = ms_all / max(1, kt["blocks"]) * 1e3
        gbps = (kt["head_alg_bytes"] + kt["tail_alg_bytes"]) / (ms_all * 1e-3) / 1e9 if ms_all > 0 else 0.0
        row = {"channels": ch, "method": "nonuniform" if method == NONUNIFORM_METHOD else "uniform",
               "p50_ms": float(np.percentile(lat, 50)), "p99_ms": p99,
               "max_ms": float(lat.max()), "tick_kernel_us": tick_us,
               "tick_GBps": gbps, "meets_deadline": bool(lat.max() <= DEADLINE_MS)}
        log(f"sweep: {row}")
        rows.append(row)
        if row["meets_deadline"]:
            best = ch
        else:
            break
    return best, rows


def bench_sweep_phased(T, torch, ctx, counts, ticks=96):
    """Non-uniform deadline sweep with the channels split over M objects of
    frame phases 0..M-1 (M: the library's frame multiple for the first
    object), advanced together in one group launch per block: each block
    carries the frame work of one object instead of all channels' frames
    falling on the same block."""
    from tan_amd import serving as S
    rows = []
    best = 0
    for ch in counts:
        free, _ = torch.cuda.mem_get_info()
        need = (S.estimate_bytes(ctx._lib, L, B, ch, method=NONUNIFORM_METHOD) + ch * L * 4
                + 2 * ch * 64 * B * 4) * 1.10
        if need > free:
            log(f"phased sweep: skip {ch} channels (needs {need / 2**30:.1f} GiB)")
            break
        first = T.TANConvolution(ctx)
        assert first.InitGpu(NONUNIFORM_METHOD, L, B, ch) == T.AMF_OK
        M = first.Info()["nu_multiple"]
        first.Terminate()
        K = max(1, M)
        bounds = [ch * k // K for k in range(K + 1)]
        objs = []
        for k in range(K):
            conv = T.TANConvolution(ctx, nu_multiple=M, nu_phase=k)
            assert conv.InitGpu(NONUNIFORM_METHOD, L, B, bounds[k + 1] - bounds[k]) == T.AMF_OK
            objs.append(conv)
        for k, conv in enumerate(objs):
            h = ir_bank_device(torch, bounds[k + 1] - bounds[k], L, 77 + k)
            torch.cuda.synchronize()
            assert conv.UpdateResponseTD(h, L, None, T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY) == 0
            del h
        x = noise_device(torch, ch, 64 * B, 5)
        y = torch.empty_like(x)
        xs = [[x[bounds[k]:bounds[k + 1], t * B:] for k in range(K)] for t in range(64)]
        ys = [[y[bounds[k]:bounds[k + 1], t * B:] for k in range(K)] for t in range(64)]
        torch.cuda.synchronize()
        for t in range(2 * M + 16):
            assert T.GroupProcess(objs, xs[t % 64], ys[t % 64]) == T.AMF_OK
        objs[0].Synchronize()
        lat = []
        for t in range(ticks):
            t0 = time.perf_counter()
            assert T.GroupProcess(objs, xs[t % 64], ys[t % 64]) == T.AMF_OK
            objs[0].Synchronize()
            lat.append((time.perf_counter() - t0) * 1e3)
        finite = bool(torch.isfinite(y).all().item())
        for conv in objs:
            conv.Terminate()
        del x, y, xs, ys
        torch.cuda.empty_cache()
        lat = np.array(lat)
        row = {"channels": ch, "method": "nonuniform-phased", "objects": K, "nu_multiple": M,
               "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
               "max_ms": float(lat.max()), "outputs_finite": finite,
               "meets_deadline": bool(lat.max() <= DEADLINE_MS)}
        log(f"phased sweep: {row}")
        rows.append(row)
        if row["meets_deadline"]:
            best = ch
        else:
            break
    return best, rows


def bench_precision(T, torch, ctx, channels=1024, taps=131072, warm=8, blocks=32):
    """Reduced-precision IR spectra on the precision workload (1,024 channels,
    B=256, 131,072 taps): block time, algorithmic GB/s, real-time channels
    and SNR of channel 0 against FP64 and against the FP32 path."""
    precs = [("fp32", T.TAN_IR_PRECISION_FP32), ("fp16", T.TAN_IR_PRECISION_FP16),
             ("bf16", T.TAN_IR_PRECISION_BF16)]
    h = ir_bank_device(torch, channels, taps, 4242)
    x = noise_device(torch, channels, (warm + blocks) * B, 4243)
    n = (warm + blocks) * B
    ref = None
    rows = []
    outs = {}
    for name, prec in precs:
        conv = T.TANConvolution(ctx, ir_precision=prec)
        assert conv.InitGpu(T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED, taps, B, channels) == T.AMF_OK
        conv.Autotune()
        torch.cuda.synchronize()
        assert conv.UpdateResponseTD(h, taps, None, T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY) == 0
        y = torch.zeros_like(x)
        conv.Process(x, y, warm * B)
        conv.Synchronize()
        conv.EnableTiming(True)
        conv.ReadTiming(reset=True)
        conv.Process(x[:, warm * B:], y[:, warm * B:], blocks * B)
        kt = conv.ReadTiming()
        conv.Terminate()
        # every launch (with multi-block passes: pass and batched head launches), per block
        ms_all = kt["head_ms"] + kt["tail_ms"]
        tick_us = ms_all / kt["blocks"] * 1e3
        gbps = (kt["head_alg_bytes"] + kt["tail_alg_bytes"]) / (ms_all * 1e-3) / 1e9
        outs[name] = y[:2].cpu().numpy().astype(np.float64)
        rows.append({"precision": name, "tick_us": tick_us, "alg_GBps": gbps, "frac": gbps / HBM_PEAK_GBPS,
                     "rt_channels": channels * DEADLINE_MS * 1e3 / tick_us})
        del y
    hx = x[:2].cpu().numpy().astype(np.float64)
    hh = h[:2].cpu().numpy().astype(np.float64)
    nfft = 1 << (n + taps - 1).bit_length()
    for row in rows:
        s64, s32 = [], []
        for c in range(2):
            full = np.fft.irfft(np.fft.rfft(hx[c], nfft) * np.fft.rfft(hh[c], nfft), nfft)[:n]
            # the IR fades in over block 2 (two blocks after the update)
            w = np.concatenate([np.zeros(2 * B), np.arange(B) / B, np.ones(n - 3 * B)])
            ref = full * w
            got = outs[row["precision"]][c]
            s64.append(10 * np.log10(np.sum(ref ** 2) / np.sum((got - ref) ** 2)))
            r32 = outs["fp32"][c]
            e32 = np.sum((got - r32) ** 2)
            s32.append(10 * np.log10(np.sum(r32 ** 2) / e32) if e32 > 0 else float("inf"))
        row["snr_vs_fp64_db"] = float(min(s64))
        row["snr_vs_fp32_db"] = float(min(s32))
        log(f"precision: {row}")
    del h, x
    torch.cuda.empty_cache()
    return rows


def _pct(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else None


def bench_batching(T, torch, ctx, counts=(512,), ticks=2000, seed=7):
    """Uneven streams (IR log-uniform 4,096-262,144 taps, B in {128, 256, 512},
    one channel each) in one group launch per 128-sample base tick, with the
    tail chunks tuned per (block, rounded length) bucket."""
    from tan_amd import batching as Bt
    rows = []
    for n in counts:
        specs = Bt.draw_streams(n, seed=seed)
        batch = Bt.StreamBatch(ctx, specs)
        period = batch.base_period_ms
        cyc = batch.cycle
        before = batch.time_ticks(cyc * 64) / (cyc * 64)
        t0 = time.perf_counter()
        chosen = batch.tune(cycles=8)
        tune_s = time.perf_counter() - t0
        after = batch.time_ticks(cyc * 64) / (cyc * 64)
        # steady stream of ticks: device time over the whole run
        total_ms = batch.time_ticks(ticks)
        per = batch.time_ticks(min(ticks, 1000), per_tick=True)
        gb = sum(batch.bytes_per_tick) / cyc / (total_ms / ticks * 1e-3) / 1e9
        row = {"streams": n, "base_block": batch.base, "deadline_ms": period,
               "objects_per_launch": [len(d) for d in batch.due],
               "tick_ms_default_chunks": before, "tick_ms_tuned": after, "tune_seconds": tune_s,
               "tuned_chunks": {f"B{b}_L{ln}": c for (b, ln), c in sorted(chosen.items())},
               "mean_tick_ms": total_ms / ticks, "p50_tick_ms": _pct(per, 50), "p99_tick_ms": _pct(per, 99),
               "max_tick_ms": float(max(per)), "deadline_misses": int(sum(1 for v in per if v > period)),
               "alg_GBps": gb, "hbm_frac": gb / HBM_PEAK_GBPS,
               "realtime_factor": period / (total_ms / ticks)}
        log(f"  batching {n} streams: tick {row['mean_tick_ms'] * 1e3:.1f} us (default chunks "
            f"{before * 1e3:.1f} us), p99 {row['p99_tick_ms'] * 1e3:.1f} us, {gb:.0f} GB/s, "
            f"tuned {row['tuned_chunks']}")
        rows.append(row)
        for c in batch.convs:
            c.Terminate()
        del batch
        torch.cuda.synchronize()
    return rows


def bench_serving(T, torch, ctx, seconds=20.0, budget_gib=32.0, seed=5, precision=None, method=UNIFORM_METHOD):
    """Poisson session arrivals (20/s, mean life 30 s, 1-16 channels, IR
    log-uniform 8,192-524,288 taps, B = 256, IR update every U{10..100}
    blocks) served with HBM-budget admission and one group launch per tick.

    The run starts from a stationary population.  Session control (admission,
    IR updates, departures) runs between ticks on the same thread; block
    latency is the device time from the tick's start marker to the end of
    its kernel (host enqueue included)."""
    from tan_amd import serving as S
    prec = T.TAN_IR_PRECISION_FP32 if precision is None else precision
    wl = S.PoissonWorkload(block=B, seed=seed)
    nb_bank = 4
    bank = torch.empty(nb_bank, 16, wl.max_len, device="cuda")
    env = torch.exp(-torch.arange(wl.max_len, device="cuda", dtype=torch.float32) / (0.25 * FS))
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    for k in range(nb_bank):
        bank[k] = torch.randn(16, wl.max_len, device="cuda", generator=g) * env
    rows_cap = 16384
    nx = 8
    noise = torch.empty(nx, rows_cap, B, device="cuda").uniform_(-1.0, 1.0, generator=g)
    torch.cuda.synchronize()
    srv = S.SessionServer(ctx, block=B, hbm_budget=int(budget_gib * (1 << 30)), ir_precision=prec,
                          io_rows=rows_cap, method=method)
    meta = {}

    def ir_for(sid, k, c, ln):
        return bank[(sid + k) % nb_bank, :c, :ln]

    def submit(d):
        meta[d["sid"]] = d
        d["updates"] = 0
        srv.submit(d["sid"], d["channels"], d["ir_len"], ir_for(d["sid"], 0, d["channels"], d["ir_len"]))

    # a stationary population: sessions that started at random earlier
    # times, admitted over a warm-up of 64 untimed ticks so that their block
    # counters (and frame phases) are spread as in steady state
    t0 = time.perf_counter()
    init = wl.initial_population()
    warm = 64
    for w in range(warm):
        for d in init[w::warm]:
            submit(d)
        srv.admit()
        srv.in_arena.copy_(noise[w % nx])
        srv.finish(srv.tick())
    setup_s = time.perf_counter() - t0
    ticks = int(seconds * FS / B)
    lat, enq, ctl, chans, sess, byts, waits = [], [], [], [], [], [], []
    admitted = departed = updates = 0
    pend = None
    # a real-time control loop: no garbage-collector pauses inside ticks
    import gc
    gc.collect()
    gc.disable()
    wall0 = time.perf_counter()
    for t in range(ticks):
        c0 = time.perf_counter()
        for sid in [sid for sid, d in meta.items() if d["end_tick"] <= t]:
            srv.release(sid)
            del meta[sid]
            departed += 1
        for d in wl.arrivals(t):
            submit(d)
        for s in srv.admit():
            admitted += 1
            waits.append(srv.ticks - s.arrived_tick)
        for sid, s in srv.active.items():
            d = meta[sid]
            if s.blocks and s.blocks % d["update_every"] == 0:
                d["updates"] += 1
                srv.update(sid, ir_for(sid, d["updates"], d["channels"], d["ir_len"]))
                updates += 1
        srv.in_arena.copy_(noise[t % nx])
        ctl.append((time.perf_counter() - c0) * 1e3)
        pend = srv.tick()
        rec = S.SessionServer.finish(pend)
        lat.append(rec.latency_ms)
        enq.append(rec.enqueue_ms)
        chans.append(rec.channels)
        sess.append(rec.sessions)
        byts.append(rec.alg_bytes)
    wall = time.perf_counter() - wall0
    gc.enable()
    finite = bool(torch.isfinite(srv.out_arena).all().item())
    queued = srv.admission.waiting()
    srv.shutdown()
    del bank, noise
    torch.cuda.synchronize()
    period = srv.period_ms
    lat_a = np.asarray(lat)
    gbps = np.asarray(byts) / (lat_a * 1e-3) / 1e9
    return {"method": "nonuniform" if method == NONUNIFORM_METHOD else "uniform",
            "block": B, "deadline_ms": period, "audio_seconds": ticks * B / FS, "ticks": ticks,
            "ir_precision": {0: "fp32", 1: "fp16", 2: "bf16"}[prec],
            "hbm_budget_GiB": budget_gib, "setup_seconds": setup_s,
            "sessions_mean": float(np.mean(sess)), "sessions_max": int(max(sess)),
            "channels_mean": float(np.mean(chans)), "channels_max": int(max(chans)),
            "admitted_during_run": admitted, "departed": departed, "ir_updates": updates,
            "queued_at_end": queued, "admission_wait_ticks_max": int(max(waits)) if waits else 0,
            "p50_block_ms": _pct(lat, 50), "p99_block_ms": _pct(lat, 99), "max_block_ms": float(lat_a.max()),
            "deadline_misses": int((lat_a > period).sum()), "miss_rate": float((lat_a > period).mean()),
            "p50_enqueue_ms": _pct(enq, 50), "p99_enqueue_ms": _pct(enq, 99),
            "p50_control_ms": _pct(ctl, 50), "p99_control_ms": _pct(ctl, 99),
            "alg_GBps_p50": _pct(gbps, 50), "hbm_frac_p50": _pct(gbps, 50) / HBM_PEAK_GBPS,
            "realtime_factor": ticks * B / FS / wall, "outputs_finite": finite}


def bench_dft_mfma(T, torch, ctx, channels=1024, log2n=9, reps=50):
    """The block DFT on the matrix cores (precision track): 1,024 channels x
    2^log2n-sample frames (256: the block; 512: the 2B-point transform at
    B = 256), plain bf16 and split-bf16, against the FP32 FFT (TANFFT R2C) on
    the same frames."""
    n = 1 << log2n
    fft = T.TANFFT(ctx)
    fft.Init()
    s = torch.cuda.ExternalStream(T._lib.load().tanx_ctx_process_stream(ctx.handle))
    g = torch.Generator(device="cuda")
    g.manual_seed(9)
    x = torch.empty(channels, 2 * n, device="cuda").uniform_(-1.0, 1.0, generator=g)
    x[:, n:] = 0.0
    torch.cuda.synchronize()
    ref = np.fft.rfft(x[:64, :n].double().cpu().numpy(), axis=1)
    refri = np.stack([ref.real, ref.imag], axis=2).reshape(64, n + 2)

    def snr(a, b):
        e = np.sum((a - b) ** 2)
        return float(10 * np.log10(np.sum(a * a) / e)) if e > 0 else float("inf")

    def timed(fn):
        fn()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3

    # the kernels alone: the library's entry points called directly (the
    # Python methods also order the library's streams after torch's, a
    # cross-queue wait per call), with torch's work finished first
    lib = T._lib.load()
    yf = torch.zeros(channels, 2 * n, device="cuda")
    torch.cuda.synchronize()
    us_fft = timed(lambda: lib.tanx_fft_transform_device(ctx.handle, T.TAN_FFT_R2C_TRANSFORM_DIRECTION_FORWARD,
                                                         log2n, channels, x.data_ptr(), yf.data_ptr(), x.stride(0)))
    torch.cuda.synchronize()
    fp32 = yf[:64, :n + 2].double().cpu().numpy()
    rows = []
    for split in (1, 3):
        y = torch.zeros(channels, 2 * n, device="cuda")
        torch.cuda.synchronize()
        us = timed(lambda: lib.tanx_dft_r2c_mfma_device(ctx.handle, log2n, channels, x.data_ptr(), x.stride(0),
                                                        y.data_ptr(), y.stride(0), split))
        torch.cuda.synchronize()
        got = y[:64, :n + 2].double().cpu().numpy()
        flops = 2.0 * channels * n * (n + 2) * split
        tf = flops / (us * 1e-6) / 1e12
        rows.append({"split": split, "channels": channels, "n": n, "us": us, "tflops": tf,
                     "peak_tflops": 2500.0, "frac": tf / 2500.0, "fft_fp32_us": us_fft,
                     "snr_vs_fp64_db": snr(refri, got), "snr_vs_fp32_fft_db": snr(fp32, got),
                     "fft_fp32_snr_vs_fp64_db": snr(refri, fp32)})
        log(f"  mfma dft n={n} split {split}: {us:.1f} us ({tf:.1f} TFLOP/s), FFT {us_fft:.1f} us, "
            f"SNR {rows[-1]['snr_vs_fp64_db']:.1f} dB vs FP64")
    del x, yf
    return rows


ROOM_SIZE = (10.0, 4.0, 8.0)  # m: width, height, length
ROOM_LEN = 65536               # response taps: the headline IR length


def _room_scene(R, count):
    """`count` sources and listeners at random places in one 10 x 4 x 8 m
    room with damping 0.9 (floor 0.6, ceiling 0.7)."""
    room = R.make_room(*ROOM_SIZE, 0.9, damp_bottom=0.6, damp_top=0.7)
    rng = np.random.default_rng(11)
    srcs = [R.MonoSource(*rng.uniform([0.5, 0.5, 0.5], [9.5, 3.5, 7.5])) for _ in range(count)]
    lsts = [R.make_listener(tuple(rng.uniform([1.0, 1.2, 1.0], [9.0, 2.0, 7.0])), 0.16,
                            float(rng.uniform(-180, 180))) for _ in range(count)]
    return room, srcs, lsts


def _room_images(length):
    n = [1 + length // int(v / 340.0 * FS) for v in ROOM_SIZE]
    return int(np.prod([2 * v + 1 for v in n]))


def bench_room(T, torch, ctx, counts=(1, 64), reps=5):
    """Room responses by image sources (IR source for the convolution):
    `count` listener/source pairs of 65,536 taps per launch, both ears."""
    from tan_amd import room as R
    s = torch.cuda.ExternalStream(T._lib.load().tanx_ctx_process_stream(ctx.handle))
    rows = []
    for count in counts:
        room, srcs, lsts = _room_scene(R, count)
        out = torch.empty((2 * count, ROOM_LEN), device="cuda")
        rooms = [room] * count
        assert R.GenerateRoomResponses(ctx, rooms, srcs, lsts, FS, ROOM_LEN, out) == 0
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(reps):
            R.GenerateRoomResponses(ctx, rooms, srcs, lsts, FS, ROOM_LEN, out)
        e1.record(s)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        images = 2 * count * _room_images(ROOM_LEN)
        row = {"responses": count, "taps": ROOM_LEN, "room_m": list(ROOM_SIZE), "ms_per_launch": ms,
               "responses_per_s": count / (ms * 1e-3), "images_per_s": images / (ms * 1e-3),
               "outputs_finite": bool(torch.isfinite(out).all().item())}
        log(f"  room responses x{count}: {ms:.3f} ms per launch, {row['responses_per_s']:.0f} responses/s, "
            f"{row['images_per_s'] / 1e9:.1f} G images/s")
        rows.append(row)
        del out
    return rows


def cpu_room_baseline():
    """The oracle's restatement of the room generator (one response, both
    ears, one thread), on the same scene."""
    from tan_amd import room as R
    path = os.path.join(ROOT, "oracle", "build", "libtan_oracle.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.orc_room_response.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_int, ctypes.c_uint,
                                                              ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    room, srcs, lsts = _room_scene(R, 1)
    left = np.zeros(ROOM_LEN, np.float32)
    right = np.zeros(ROOM_LEN, np.float32)
    t0 = time.perf_counter()
    rc = lib.orc_room_response(ctypes.byref(room), ctypes.byref(srcs[0]), ctypes.byref(lsts[0]), FS, ROOM_LEN, 0, 0,
                               left.ctypes.data, right.ctypes.data)
    dt = time.perf_counter() - t0
    return {"responses_per_s": 1.0 / dt, "cores": 1, "kind": "port", "rc": rc,
            "sample": f"one {ROOM_LEN}-tap response (both ears) in the same room, scalar C"}


def pmc_traffic(grid_threads, kernel_prefix="void tanx::k_tick<256, 1, 0, 1>"):
    """HBM bytes per launch from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE scaled by a 512 MiB copy), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic_k_tick.json")
    try:
        data = json.load(open(path))
    except (OSError, ValueError):
        return None
    for k in data.get("kernels", []):
        if k["kernel"].startswith(kernel_prefix) and k["grid_threads"] == grid_threads:
            return float(k["traffic_bytes"])
    return None


def cpu_baseline(seconds):
    """The oracle's restatement of the reference CPU path, timed on this host."""
    path = os.path.join(ROOT, "oracle", "build", "libtan_oracle.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.orc_bench_upols.restype = ctypes.c_double
    lib.orc_bench_upols.argtypes = [ctypes.c_int] * 5 + [ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    procs = lib.orc_num_procs()
    # reference thread policy: omp_get_num_procs()/4 (FFTImpl.cpp:245-250)
    threads = lib.orc_set_threads(max(1, procs // 4))
    probe = 8
    t = lib.orc_bench_upols(C, L, B, probe, 4, 11, None)
    per_block = t / probe
    nblocks = int(max(16, min(20000, seconds / max(per_block, 1e-6))))
    times = (ctypes.c_double * nblocks)()
    total = lib.orc_bench_upols(C, L, B, nblocks, 8, 12, times)
    arr = np.array(times[:])
    rt = C * nblocks * B / FS / total
    return {"value": rt, "unit": "real-time channels", "cores": int(threads), "kind": "port",
            "sample": (f"{nblocks} blocks of the same 64ch/B=256/65536-tap workload (one IR, no "
                       f"updates), {threads} OpenMP threads of {procs} procs (reference policy "
                       f"procs/4), own radix-2 FFT (FFTW absent)"),
            "p99_block_ms": float(np.percentile(arr, 99) * 1e3),
            "mean_block_ms": float(arr.mean() * 1e3)}


def profile_run(T, torch, ctx):
    """Short fixed workload for counter collection (rocprofv3 --pmc): 100
    blocks of the headline configuration, 10 blocks at 8,192 channels
    (HBM-bound), and a plain device copy of 512 MiB (3 launches) whose
    FETCH_SIZE/WRITE_SIZE calibrate the counters against known bytes."""
    bench_throughput(T, torch, ctx, 100, 10, timing=False)
    # the headline shape at every tail chunk count the autotuner may pick
    conv = T.TANConvolution(ctx)
    assert conv.InitGpu(T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED, L, B, C) == T.AMF_OK
    h = ir_bank_device(torch, C, L, 78)
    torch.cuda.synchronize()
    conv.UpdateResponseTD(h, L, None, T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY)
    x = noise_device(torch, C, 20 * B, 6)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for chunks in (1, 2, 4, 8, 16, 32, 64):
        conv.SetTailChunks(chunks)
        conv.Process(x, y, 20 * B)
    conv.Synchronize()
    conv.Terminate()
    del h, x, y
    ch = 8192
    conv = T.TANConvolution(ctx)
    assert conv.InitGpu(T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED, L, B, ch) == T.AMF_OK
    h = ir_bank_device(torch, ch, L, 77)
    torch.cuda.synchronize()
    conv.UpdateResponseTD(h, L, None, T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY)
    del h
    x = noise_device(torch, ch, 16 * B, 5)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    conv.Process(x, y, 10 * B)
    conv.Synchronize()
    conv.Terminate()
    del x, y
    a = torch.ones(512 << 18, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("profile run done")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--latency-blocks", type=int, default=300)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tail-chunks", type=int, default=0,
                    help="pin the throughput passes' tail chunks instead of autotuning (for profiled runs)")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-batching", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--batching-streams", type=str, default="512,2048")
    ap.add_argument("--serving-seconds", type=float, default=15.0)
    ap.add_argument("--no-nonuniform", action="store_true")
    ap.add_argument("--no-room", action="store_true")
    ap.add_argument("--serving-methods", type=str, default="uniform,nonuniform")
    ap.add_argument("--nu-sweep", type=str, default="32768,65536,98304,131072")
    ap.add_argument("--sweep", type=str, default="1024,4096,8192,16384,24576,28672,32768,36864,40960")
    ap.add_argument("--profile-run", action="store_true",
                    help="short fixed workload for rocprofv3 counter passes; prints no JSON")
    args = ap.parse_args()
    global TAIL_CHUNKS
    TAIL_CHUNKS = args.tail_chunks
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (channels and "
              "sessions shard across GPUs without exchange; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import torch
    import tan_amd as T
    T.load_library()
    ctx = T.TANContext(device=0)
    # The throughput passes use a context created on a torch stream (the
    # caller-stream form of the reference's InitOpenCL(queue, queue)): the
    # inputs are written on that stream, so Process needs no cross-stream
    # wait per call, as a C++ caller on its own stream would not.
    hs = torch.cuda.Stream()
    hctx = T.TANContext(device=0, conv_stream=hs, general_stream=hs)
    if args.profile_run:
        profile_run(T, torch, ctx)
        hctx.Terminate()
        ctx.Terminate()
        return

    log("throughput pass")
    dt, _, info, finite = bench_throughput(T, torch, hctx, args.steps, args.warmup, timing=False, stream=hs)
    assert finite, "non-finite output"
    value = C * args.steps * B / FS / dt
    log(f"  {args.steps} blocks in {dt * 1e3:.1f} ms -> {value:.0f} real-time channels")

    log("kernel timing pass")
    dt2, kt, info2, _ = bench_throughput(T, torch, hctx, args.steps, args.warmup, timing=True, stream=hs)
    kp = info2.get("pass_blocks", 1)
    head_blocks = None
    if kp > 1:
        # a launch runs up to kp heads and the pass computing the next kp
        # blocks' partial sums (the dominant kernel, k_window<256, 1, 0>;
        # k_tick<256, 1, 0, 2> where a launch has a single head); launches
        # with heads alone are timed as head launches
        tick_us = kt["tail_ms"] / kt["tail_launches"] * 1e3
        achieved = kt["tail_alg_bytes"] / (kt["tail_ms"] * 1e-3) / 1e9
        alg_per_launch = kt["tail_alg_bytes"] / kt["tail_launches"]
        launches = kt["tail_launches"]
        kernel = f"k_window (heads + {kp}-block pass)"
        head_only_us = kt["head_ms"] / kt["head_launches"] * 1e3 if kt["head_launches"] else None
        head_blocks = kt["blocks"] / (kt["head_launches"] + kt["tail_launches"])
        prefix = "void tanx::k_window<256, 1, 0>"
    else:
        # one k_tick launch per block: the block's head and the next block's tail
        tick_us = kt["head_ms"] / kt["head_launches"] * 1e3
        achieved = kt["head_alg_bytes"] / (kt["head_ms"] * 1e-3) / 1e9
        alg_per_launch = kt["head_alg_bytes"] / kt["head_launches"]
        launches = kt["head_launches"]
        kernel = "k_tick (head + next tail)"
        head_only_us = None
        prefix = "void tanx::k_tick<256, 1, 0, 1>"
    per_block_us = (kt["head_ms"] + kt["tail_ms"]) * 1e3 / kt["blocks"]
    # the launch's grid: 64 head blocks + 64 x chunks tail blocks of 256 threads
    traffic = pmc_traffic((C + C * info2["tail_chunks"]) * 256, prefix)
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic, "kernel": kernel,
                "traffic_source": "profiles/pmc_traffic_k_tick.json" if traffic else None,
                "alg_bytes_per_launch": alg_per_launch, "avg_launch_us": tick_us, "launches": launches,
                "pass_blocks": kp, "head_launch_us": head_only_us, "blocks_per_launch": head_blocks,
                "avg_block_us": per_block_us,
                "note": ("64-channel working set (~67 MB) is resident in the 256 MB Infinity "
                         "Cache; see sweep for HBM-bound channel counts")}
    log(f"  tick kernel {tick_us:.2f} us/launch, {achieved:.0f} GB/s algorithmic; "
        f"{per_block_us:.2f} us per block on average")

    log("latency pass")
    lat = bench_latency(T, ctx, args.latency_blocks)
    log(f"  {lat}")

    best, rows = (0, [])
    if not args.no_sweep:
        log("deadline sweep")
        best, rows = bench_sweep(T, torch, ctx, [int(v) for v in args.sweep.split(",") if v])

    prec_rows, dft_rows = [], []
    if not args.no_precision:
        log("precision pass")
        prec_rows = bench_precision(T, torch, ctx)
        # 256 points (the configuration's block DFT) and 512 (the head's 2B-point frame)
        dft_rows = bench_dft_mfma(T, torch, ctx, log2n=8) + bench_dft_mfma(T, torch, ctx, log2n=9)

    nonuniform = None
    if not args.no_nonuniform:
        log("non-uniform pass")
        dtn, _, infon, finite_n = bench_throughput(T, torch, hctx, args.steps, args.warmup, timing=False,
                                                   method=NONUNIFORM_METHOD, stream=hs)
        assert finite_n, "non-finite output (non-uniform)"
        nu_value = C * args.steps * B / FS / dtn
        log(f"  non-uniform (M = {infon['nu_multiple']}): {nu_value:.0f} real-time channels")
        nu_best, nu_rows = (0, [])
        ph_best, ph_rows = (0, [])
        if not args.no_sweep:
            counts = [int(v) for v in args.nu_sweep.split(",") if v]
            nu_best, nu_rows = bench_sweep(T, torch, ctx, counts, method=NONUNIFORM_METHOD)
            ph_best, ph_rows = bench_sweep_phased(T, torch, ctx, counts)
        nonuniform = {"method": "TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM",
                      "nu_multiple": infon["nu_multiple"], "value": nu_value,
                      "ms_per_step": dtn / args.steps * 1e3, "rt_channels_at_deadline": nu_best,
                      "sweep": nu_rows, "rt_channels_at_deadline_phased": ph_best,
                      "sweep_phased": ph_rows}

    room_rows = []
    if not args.no_room:
        log("room response pass")
        room_rows = bench_room(T, torch, ctx)

    batching = []
    if not args.no_batching:
        log("batching pass")
        batching = bench_batching(T, torch, ctx, [int(v) for v in args.batching_streams.split(",") if v])

    serving = []
    if not args.no_serving:
        names = {"uniform": UNIFORM_METHOD, "nonuniform": NONUNIFORM_METHOD}
        for meth in [names[m] for m in args.serving_methods.split(",") if m]:
            log(f"serving pass ({'non-uniform' if meth == NONUNIFORM_METHOD else 'uniform'})")
            row = bench_serving(T, torch, ctx, seconds=args.serving_seconds, method=meth)
            log(f"  {row}")
            serving.append(row)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(args.cpu_seconds)
        log(f"  {cpu}")
        if room_rows:
            cpu_room = cpu_room_baseline()
            log(f"  room generator: {cpu_room}")
            for r in room_rows:
                r["cpu_responses_per_s"] = cpu_room["responses_per_s"] if cpu_room else None
    hctx.Terminate()
    ctx.Terminate()

    big = max(rows, key=lambda r: r["channels"]) if rows else None
    out = {
        "metric": METRIC, "value": value, "unit": "real-time channels", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (decaying-noise IRs, U[-1,1] input, generated on device)",
        "config": {"workload": WORKLOAD, "channels": C, "block": B, "ir_taps": L,
                   "blocks": args.steps, "update_every": UPDATE_EVERY, "sample_rate": FS},
        "p50_block_ms": lat["p50_ms"], "p99_block_ms": lat["p99_ms"], "max_block_ms": lat["max_ms"],
        "deadline_ms": DEADLINE_MS, "deadline_misses": lat["deadline_misses"],
        "latency_worst_blocks": lat["worst"],
        "roofline": roofline,
        "rt_channels_at_deadline": best,
        "sweep": rows,
        "roofline_hbm": ({"channels": big["channels"], "achieved": big["tick_GBps"],
                          "frac": big["tick_GBps"] / HBM_PEAK_GBPS} if big else None),
        "precision": prec_rows,
        "mfma_dft": dft_rows,
        "nonuniform": nonuniform,
        "batching": batching,
        "serving": serving,
        "room_responses": room_rows,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": (value / cpu["value"]) if cpu else None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
