#!/usr/bin/env python3
"""Benchmark of the time-varying multichannel convolution path on one MI355X.

value: the real-time channel count at a 65,536-tap IR and 256-sample blocks
(48 kHz): the most channels for which every one of `--steps` x 100
consecutive ticks finishes within the 5.333 ms block period, one block in
flight, with a fresh IR for every channel every 50 blocks and a one-block
crossfade.  Channels are grouped in objects of 2,048 whose multi-block
passes run on staggered blocks (pass phases) and whose IR updates are spread
over the 50-block period; all groups advance in one group launch per tick.
Inputs, outputs and IR banks are resident in HBM before the timed region.
A step is 100 ticks (two update periods), so the default 20 steps are 2,000
ticks.

Also reported (one JSON line):
  roofline      -- the dominant kernel of the timed run (the group tick
                   kernel): algorithmic bytes of the roles it ran / its
                   launch-stamped duration, against 8 TB/s, with the HBM
                   traffic per launch from the committed PMC summary;
  c2_64ch       -- the 64-channel workload as configured (2,000 blocks, IR
                   updates every 50 inside the timed window);
  latency_64ch  -- one block per call with host buffers (PCIe included);
  precision, mfma_dft, nonuniform, batching, serving, room_responses;
  cpu_baseline  -- the oracle's restatement of the reference CPU path on a
                   bounded sample of the 64-channel workload, with the
                   reference's thread policy and with all cores.
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
RT_WORKLOAD = ("real-time channel count of the plain FP32 method: IR 65,536 taps (decaying noise), block 256 "
               "at 48 kHz, a fresh IR for every channel every 50 blocks with a one-block linear crossfade, "
               "every tick (one block in flight, device-resident data) within 5.333 ms over steps x 100 "
               "consecutive ticks; channels in groups of 2,048 with staggered passes and updates, one group "
               "launch per tick")


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


RT_GROUP = 2048        # channels per object (channel group) of the real-time count
RT_STEP_TICKS = 100    # ticks per timed step: two IR update periods
RT_RING = 8            # input/output blocks kept on the device
RT_BANKS = 4           # IR banks the updates cycle through
RT_PASS = 12           # blocks per tail pass of the real-time groups (phases rt_phase(g))
RT_SLOTS = 2           # IR spectrum slots per object (2: a third less HBM per channel; the
                       # schedule never updates a group while its switch is armed)


def rt_chunks():
    """Partition chunks per channel of a real-time group's pass: the
    library's default for the group size (one launch carries every group,
    so wider passes buy nothing), which also sizes the partial rows kept."""
    return max(1, 1024 // RT_GROUP)


def rt_balanced():
    """Groups of RT_GROUP channels per pass phase before the smaller ones
    (72 large groups in all at 8- or 12-block passes, 64 at 16)."""
    return max(1, 72 // RT_PASS)


def rt_large():
    """Groups of RT_GROUP channels: rt_balanced() per pass phase; the groups
    after them have half as many channels, one per phase in turn, so a
    count between two multiples of RT_PASS large groups adds half a large
    group's pass to some phases instead of a whole one."""
    return rt_balanced() * RT_PASS


def rt_group_channels(g):
    """Channels of group g."""
    return RT_GROUP if g < rt_large() else RT_GROUP // 2


def rt_channels(k):
    """Channels of the first k groups."""
    return sum(rt_group_channels(g) for g in range(k))


def rt_phase(g):
    """Pass phase of group g: phases in the order 0, 2, 4, ..., 1, 3, 5, ...
    round robin over the large groups, then again over the smaller ones, so
    the phases that carry one group more than others pass on even ticks."""
    order = list(range(0, RT_PASS, 2)) + list(range(1, RT_PASS, 2))
    return order[(g if g < rt_large() else g - rt_large()) % RT_PASS]


def rt_update_offset(g):
    """Update offset of group g in the 50-tick period: one group per offset
    for the first 50, then the odd offsets 1, 3, 5, ... for the other large
    groups, so that ticks with two large updates are odd ticks; the smaller
    groups take the even offsets (ticks with one large update), then the
    odd offsets the large groups left."""
    if g < UPDATE_EVERY:
        return g
    big = rt_large()
    if g < big:
        return (2 * (g - UPDATE_EVERY) + 1) % UPDATE_EVERY
    used = {(2 * (j - UPDATE_EVERY) + 1) % UPDATE_EVERY for j in range(UPDATE_EVERY, big)}
    seq = list(range(0, UPDATE_EVERY, 2)) + [o for o in range(1, UPDATE_EVERY, 2) if o not in used]
    return seq[(g - big) % len(seq)]


class RealTimeRig:
    """The real-time channel count: channel groups of rt_group_channels(g)
    channels (one object each, 65,536 taps, B = 256, RT_PASS-block passes
    with phases rt_phase(g)), a fresh IR for every channel every 50 blocks (group g's
    update at the ticks t with (t + rt_update_offset(g)) % 50 == 0, so
    updates and their crossfades are staggered over the period), all k
    active groups advanced by one group launch per tick with one block in
    flight.  Everything runs on one torch stream that is the context's
    process stream."""

    def __init__(self, T, torch, ngroups):
        self.T, self.torch = T, torch
        self.hs = torch.cuda.Stream()
        self.ctx = T.TANContext(device=0, conv_stream=self.hs, general_stream=self.hs)
        self.k = ngroups
        self.t = 0
        self.off = [rt_channels(g) for g in range(ngroups + 1)]  # first channel of group g
        with torch.cuda.stream(self.hs):
            self.banks = [ir_bank_device(torch, RT_GROUP, L, 5000 + i) for i in range(RT_BANKS)]
            total = self.off[ngroups]
            self.x = noise_device(torch, RT_RING * total, B, 11).view(RT_RING, total, B)
            self.y = torch.zeros_like(self.x)
            self.objs = []
            for g in range(ngroups):
                conv = T.TANConvolution(self.ctx, pass_blocks=RT_PASS, pass_phase=rt_phase(g),
                                        max_tail_chunks=rt_chunks(), ir_slots=RT_SLOTS)
                rc = conv.InitGpu(UNIFORM_METHOD, L, B, rt_group_channels(g))
                if rc != T.AMF_OK:
                    log(f"  group {g}: InitGpu returned {rc}; {g} groups")
                    break
                # no IR yet: each group's first comes from its own update
                # schedule (within the warm-up), so every update is 50 blocks
                # after the previous one (two IR slots refuse one that finds
                # a switch still armed)
                self.objs.append(conv)
        self.kmax = len(self.objs)
        self.updates = [0] * self.kmax
        self.calls = {}
        self.ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(2)]
        torch.cuda.synchronize()

    def _call(self, k, r):
        from tan_amd.batching import GroupCall
        key = (k, r)
        if key not in self.calls:
            rows = lambda a, g: a[r, self.off[g]:self.off[g + 1]]  # noqa: E731
            self.calls[key] = GroupCall(self.objs[:k], [rows(self.x, g) for g in range(k)],
                                        [rows(self.y, g) for g in range(k)])
        return self.calls[key]

    def tick(self, k):
        """One tick of the first k groups: their due IR updates, then one group
        launch; returns (host ms, device ms) from the tick's start to the end
        of its kernels (one block in flight)."""
        T = self.T
        e0, e1 = self.ev[self.t & 1]
        t0 = time.perf_counter()
        e0.record(self.hs)
        due, irs = [], []
        for g in range(k):
            # fixed offsets (a group keeps its 50-block period while the count
            # of active groups changes)
            if (self.t + rt_update_offset(g)) % UPDATE_EVERY == 0:
                self.updates[g] += 1
                due.append(self.objs[g])
                irs.append(self.banks[(self.updates[g] + g) % RT_BANKS][:rt_group_channels(g)])
        if due:
            # the tick's updates in one call: their IR transforms run in one
            # launch on the process stream, ahead of this tick's group launch
            rc = T.GroupUpdateResponseTD(due, irs, [L] * len(due))
            assert rc == T.AMF_OK, f"grouped IR update of {len(due)} groups: AMF_RESULT {rc}"
        rc = self._call(k, self.t % RT_RING).launch(order=False)
        assert rc == T.AMF_OK, rc
        e1.record(self.hs)
        e1.synchronize()
        host = (time.perf_counter() - t0) * 1e3
        self.t += 1
        return host, e0.elapsed_time(e1)

    def run(self, k, n):
        hosts, devs = [], []
        with self.torch.cuda.stream(self.hs):
            for _ in range(n):
                h, d = self.tick(k)
                hosts.append(h)
                devs.append(d)
        return np.asarray(hosts), np.asarray(devs)

    def close(self):
        for o in self.objs:
            o.Terminate()
        self.ctx.Terminate()
        self.objs, self.calls, self.banks = [], {}, []
        self.x = self.y = None
        self.torch.cuda.synchronize()
        self.torch.cuda.empty_cache()


def bench_realtime(T, torch, steps, warmup, margin=0.98):
    """Largest channel count whose every tick of `steps` x 100 consecutive
    ticks (IR updates and crossfades included) finishes within the 5.333 ms
    block period with one block in flight.

    Groups are added while HBM holds them; probes then lower the number of
    active groups from that maximum (never raising it, so the active groups
    keep equal block counts and their pass phases stay staggered) until a
    probe of 100 ticks stays under margin x deadline; the timed run of
    exactly `steps` steps follows, repeated with 2% fewer groups (at least
    one) after any miss.  The count is that of a run without a miss."""
    from tan_amd import serving as S
    gc_ = __import__("gc")
    sizes = {}

    def group_bytes(g):
        c = rt_group_channels(g)
        if c not in sizes:
            sizes[c] = S.estimate_bytes(T._lib.load(), L, B, c, pass_blocks=RT_PASS, max_tail_chunks=rt_chunks(),
                                        ir_slots=RT_SLOTS) + 2 * RT_RING * c * B * 4
        return sizes[c]

    free, _ = torch.cuda.mem_get_info()
    reserve = RT_BANKS * RT_GROUP * L * 4 + (10 << 30)
    kmax, used = 0, 0
    while used + group_bytes(kmax) <= free - reserve:
        used += group_bytes(kmax)
        kmax += 1
    log(f"real-time count: {free / 2**30:.0f} GiB free, {group_bytes(0) / 2**20:.0f} MiB per group of {RT_GROUP} "
        f"channels ({rt_balanced()} per pass phase), then groups of {RT_GROUP // 2}: up to {kmax} groups, "
        f"{rt_channels(kmax)} channels")
    rig = RealTimeRig(T, torch, kmax)
    try:
        k = rig.kmax
        # a real-time loop: no garbage-collector pauses inside ticks
        gc_.collect()
        gc_.disable()
        rig.run(k, 2 * UPDATE_EVERY)  # warm-up: every group updated and crossfaded once
        probes = []
        while k > 0:
            h, d = rig.run(k, 2 * UPDATE_EVERY)
            p99 = float(np.percentile(h, 99))
            probes.append({"groups": k, "channels": rt_channels(k), "max_host_ms": float(h.max()),
                           "p99_host_ms": p99, "max_device_ms": float(d.max())})
            log(f"  probe {probes[-1]}")
            if h.max() <= margin * DEADLINE_MS:
                break
            # tick time grows with the active groups, less than in proportion
            # (a tick has a fixed part): step half way to the count the p99
            # tick scales to (a lone outlier does not steer the search), at
            # least one group and at most 30% down per probe
            k_lin = int(k * margin * DEADLINE_MS / p99)
            k = max(1, max(int(k * 0.7), min(k - 1, k - -(-(k - k_lin) // 2))))
        gc_.enable()
        result = None
        for attempt in range(12):
            rig.run(k, warmup * RT_STEP_TICKS)
            gc_.collect()
            gc_.disable()
            rig.ctx.EnableGroupTiming(True)
            # a few ticks after the pauses above (collection, set-up), as in a
            # loop that keeps a tick every 5.333 ms: the first tick after the
            # device idles runs slower while it returns to full clock
            rig.run(k, 8)
            rig.ctx.ReadGroupTiming(reset=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            h, d = rig.run(k, steps * RT_STEP_TICKS)
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            gc_.enable()
            gt = rig.ctx.ReadGroupTiming(reset=True)
            rig.ctx.EnableGroupTiming(False)
            misses = int((h > DEADLINE_MS).sum())
            result = {"groups": k, "channels": rt_channels(k), "ticks": int(h.size), "wall_s": wall,
                      "misses": misses, "p50_host_ms": float(np.percentile(h, 50)),
                      "p99_host_ms": float(np.percentile(h, 99)), "max_host_ms": float(h.max()),
                      "p50_device_ms": float(np.percentile(d, 50)), "p99_device_ms": float(np.percentile(d, 99)),
                      "max_device_ms": float(d.max()), "max_tick_index": int(np.argmax(h)),
                      "group_timing": gt, "attempt": attempt,
                      "probes": probes}
            log(f"  timed run: {k} groups, {misses} misses, max {h.max():.3f} ms (tick {int(np.argmax(h))}), "
                f"p99 {result['p99_host_ms']:.3f} ms")
            if misses == 0 or k == 1:
                break
            k = max(1, k - max(1, -(-k // 50)))
        if result["misses"]:
            # no run without a miss: no real-time count
            result["channels"] = 0
        # in slices: HBM may be nearly full with the groups' state
        yv = rig.y.view(-1)
        finite = all(bool(torch.isfinite(yv[i:i + (1 << 24)]).all().item()) for i in range(0, yv.numel(), 1 << 24))
        result["outputs_finite"] = finite
        return result
    finally:
        rig.close()


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
    partition layout (uniform with 4- or 8-block passes, or non-uniform
    frames of M x B) and then the tail chunks tuned per (block, rounded
    length) bucket, by measured tick time."""
    from tan_amd import batching as Bt
    rows = []
    for n in counts:
        specs = Bt.draw_streams(n, seed=seed)
        batch = Bt.StreamBatch(ctx, specs)
        period = batch.base_period_ms
        cyc = batch.cycle
        batch.time_ticks(cyc * 16)  # past the first passes and the IRs' fade-in
        before = batch.time_ticks(cyc * 64) / (cyc * 64)
        t0 = time.perf_counter()
        layouts = batch.tune_layouts(cycles=8)
        after_layout = batch.time_ticks(cyc * 64) / (cyc * 64)
        chosen = batch.tune(cycles=8)
        after = batch.time_ticks(cyc * 64) / (cyc * 64)
        kept_default = False
        if after > before:
            # the tuned batch measured slower than the library's choices
            # (ticks near their host enqueue time): keep the defaults
            batch.restore_defaults()
            batch.time_ticks(cyc * 16)  # rebuilt from block 0: past the first passes again
            after = batch.time_ticks(cyc * 64) / (cyc * 64)
            layouts = {k: None for k in layouts}
            chosen = {k: 0 for k in chosen}
            kept_default = True
        tune_s = time.perf_counter() - t0
        # the same ticks as work-queue launches (the device's resident
        # workgroups take each launch's items from a counter)
        ctx.SetGroupQueue(True)
        batch.time_ticks(cyc * 8)
        queue_ms = batch.time_ticks(cyc * 64) / (cyc * 64)
        ctx.SetGroupQueue(False)
        batch.time_ticks(cyc * 8)
        # steady stream of ticks: device time over the whole run, and the
        # launch-stamped group kernels with the bytes their roles streamed
        ctx.EnableGroupTiming(True)
        ctx.ReadGroupTiming(reset=True)
        total_ms = batch.time_ticks(ticks)
        gt = ctx.ReadGroupTiming(reset=True)
        ctx.EnableGroupTiming(False)
        per = batch.time_ticks(min(ticks, 1000), per_tick=True)
        kms = gt["tick_ms"] + gt["frame_ms"]
        gb = (gt["tick_alg_bytes"] + gt["frame_alg_bytes"]) / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
        row = {"streams": n, "base_block": batch.base, "deadline_ms": period,
               "objects_per_launch": [len(d) for d in batch.due],
               "tick_ms_default": before, "tick_ms_layouts_tuned": after_layout, "tick_ms_tuned": after,
               "tick_ms_queue": queue_ms,
               "tune_seconds": tune_s,
               "kept_default": kept_default,
               "tuned_layouts": {f"B{b}_L{ln}": (f"{v[0]} {v[1]}" if v else "default")
                                 for (b, ln), v in sorted(layouts.items())},
               "tuned_chunks": {f"B{b}_L{ln}": c for (b, ln), c in sorted(chosen.items())},
               "mean_tick_ms": total_ms / ticks, "p50_tick_ms": _pct(per, 50), "p99_tick_ms": _pct(per, 99),
               "max_tick_ms": float(max(per)), "deadline_misses": int(sum(1 for v in per if v > period)),
               "kernel_alg_GBps": gb, "hbm_frac": gb / HBM_PEAK_GBPS,
               "kernel_ms_per_tick": kms / ticks, "alg_MB_per_tick": (gt["tick_alg_bytes"] + gt["frame_alg_bytes"])
               / ticks / 1e6, "realtime_factor": period / (total_ms / ticks)}
        log(f"  batching {n} streams: tick {row['mean_tick_ms'] * 1e3:.1f} us (default "
            f"{before * 1e3:.1f} us, layouts tuned {after_layout * 1e3:.1f} us, tuned {after * 1e3:.1f} us"
            f"{', defaults kept' if kept_default else ''}, work-queue launches {queue_ms * 1e3:.1f} us), p99 "
            f"{row['p99_tick_ms'] * 1e3:.1f} us, {gb:.0f} GB/s, layouts {row['tuned_layouts']}")
        rows.append(row)
        for c in batch.convs:
            c.Terminate()
        del batch
        torch.cuda.synchronize()
    return rows


def bench_serving(T, torch, ctx, seconds=20.0, budget_gib=96.0, seed=5, precision=None, method=UNIFORM_METHOD,
                  pass_blocks=8):
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
                          io_rows=rows_cap, method=method, pass_blocks=pass_blocks)
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
    lat, enq, ctl, chans, sess, waits = [], [], [], [], [], []
    admitted = departed = updates = 0
    pend = None
    # a real-time control loop: no garbage-collector pauses inside ticks
    import gc
    gc.collect()
    gc.disable()
    # launch-stamped group kernels and the bytes their roles streamed
    ctx.EnableGroupTiming(True)
    ctx.ReadGroupTiming(reset=True)
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
    wall = time.perf_counter() - wall0
    gc.enable()
    gt = ctx.ReadGroupTiming(reset=True)
    ctx.EnableGroupTiming(False)
    finite = bool(torch.isfinite(srv.out_arena).all().item())
    queued = srv.admission.waiting()
    srv.shutdown()
    del bank, noise
    torch.cuda.synchronize()
    period = srv.period_ms
    lat_a = np.asarray(lat)
    kms = gt["tick_ms"] + gt["frame_ms"]
    gbps = (gt["tick_alg_bytes"] + gt["frame_alg_bytes"]) / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    return {"method": "nonuniform" if method == NONUNIFORM_METHOD else "uniform",
            "block": B, "deadline_ms": period, "audio_seconds": ticks * B / FS, "ticks": ticks,
            "ir_precision": {0: "fp32", 1: "fp16", 2: "bf16"}[prec],
            "hbm_budget_GiB": budget_gib, "setup_seconds": setup_s,
            "pass_blocks": pass_blocks if method == UNIFORM_METHOD else None,
            "pass_phases": "balanced at admission" if method == UNIFORM_METHOD else None,
            "ir_updates_per_tick_call": "one grouped call (tanx_group_update_td_device)",
            "sessions_mean": float(np.mean(sess)), "sessions_max": int(max(sess)),
            "channels_mean": float(np.mean(chans)), "channels_max": int(max(chans)),
            "admitted_during_run": admitted, "departed": departed, "ir_updates": updates,
            "queued_at_end": queued, "admission_wait_ticks_max": int(max(waits)) if waits else 0,
            "p50_block_ms": _pct(lat, 50), "p99_block_ms": _pct(lat, 99), "max_block_ms": float(lat_a.max()),
            "deadline_misses": int((lat_a > period).sum()), "miss_rate": float((lat_a > period).mean()),
            "p50_enqueue_ms": _pct(enq, 50), "p99_enqueue_ms": _pct(enq, 99),
            "p50_control_ms": _pct(ctl, 50), "p99_control_ms": _pct(ctl, 99),
            "kernel_alg_GBps": gbps, "hbm_frac": gbps / HBM_PEAK_GBPS,
            "kernel_ms_per_tick": kms / ticks,
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


def cpu_model():
    """The host CPU's model name and logical CPU count (from /proc/cpuinfo)."""
    name, n = None, 0
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                n += 1
                name = name or line.split(":", 1)[1].strip()
    except OSError:
        pass
    return name, n


def cpu_baseline(seconds):
    """The oracle's restatement of the reference CPU path (UPOLS with an FDL,
    planar FMA multiply-accumulate, FP32), timed on this host on a bounded
    sample of the 64-channel workload: with the reference's thread policy
    (OpenMP threads = procs / 4, FFTImpl.cpp:245-250) as the reported value,
    and with every core as a stronger baseline beside it."""
    path = os.path.join(ROOT, "oracle", "build", "libtan_oracle.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    lib.orc_bench_upols.restype = ctypes.c_double
    lib.orc_bench_upols.argtypes = [ctypes.c_int] * 5 + [ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]
    lib.orc_set_threads.argtypes = [ctypes.c_int]
    procs = lib.orc_num_procs()
    model, logical = cpu_model()

    def run(nthreads, budget):
        threads = lib.orc_set_threads(nthreads)
        probe = 8
        t = lib.orc_bench_upols(C, L, B, probe, 4, 11, None)
        nblocks = int(max(16, min(20000, budget / max(t / probe, 1e-6))))
        times = (ctypes.c_double * nblocks)()
        total = lib.orc_bench_upols(C, L, B, nblocks, 8, 12, times)
        arr = np.array(times[:])
        return threads, nblocks, C * nblocks * B / FS / total, arr

    th, nb, rt, arr = run(max(1, procs // 4), seconds)
    tha, nba, rta, arra = run(procs, seconds / 2)
    fft = "the oracle's radix-2 FFT (FFTW 3.3.5 and MKL absent)"
    return {"value": rt, "unit": "real-time channels", "cores": int(th), "kind": "port",
            "sample": (f"{nb} blocks of the 64-channel, B=256, 65,536-tap workload (one IR, no updates), "
                       f"{th} OpenMP threads of {procs} processors (reference policy procs/4), {fft}"),
            "p99_block_ms": float(np.percentile(arr, 99) * 1e3), "mean_block_ms": float(arr.mean() * 1e3),
            "cpu_model": model, "logical_cpus": logical, "omp_procs": procs, "fft": fft,
            "all_cores": {"value": rta, "threads": int(tha), "blocks": nba,
                          "p99_block_ms": float(np.percentile(arra, 99) * 1e3),
                          "mean_block_ms": float(arra.mean() * 1e3)}}


def profile_run(T, torch, ngroups):
    """Short fixed workload for counter collection (rocprofv3 --pmc): the
    real-time rig with `ngroups` groups of 2,048 channels for 200 ticks (its
    updates and crossfades included), and a plain device copy of 512 MiB (3
    launches) whose FETCH_SIZE / WRITE_SIZE calibrate the counters against
    known bytes."""
    rig = RealTimeRig(T, torch, ngroups)
    h, d = rig.run(rig.kmax, 200)
    log(f"profile run: {rig.kmax} groups ({rt_channels(rig.kmax)} channels), 200 ticks, max {h.max():.3f} ms")
    rig.close()
    a = torch.ones(512 << 18, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("profile run done")


PROFILE_CMDS = {
    "trace": "rocprofv3 --kernel-trace --stats --output-format csv -d <repo>/gpurun_out/prof_rt -o run -- "
             "python3 <repo>/bench.py --profile-run --rt-groups {g}",
    "fetch": "rocprofv3 --pmc FETCH_SIZE --output-format csv -d <repo>/gpurun_out/pmc_fetch -o run -- "
             "python3 <repo>/bench.py --profile-run --rt-groups {g}",
    "write": "rocprofv3 --pmc WRITE_SIZE --output-format csv -d <repo>/gpurun_out/pmc_write -o run -- "
             "python3 <repo>/bench.py --profile-run --rt-groups {g}",
}
PROFILE_TICKS = 200
PROFILE_GROUPS = 72    # groups of the profiled rig runs (about the real-time count)
PMC_JSON = os.path.join(ROOT, "profiles", "pmc_traffic_k_group_tick_rt.json")


PROFILE_LEGS = {
    "c2": ("k_window_64ch_65536taps_b256.txt",
           "one 64-channel object x 65,536 taps, B = 256, 2,000 blocks in calls of 50 blocks, a fresh IR "
           "every 50 blocks (the c2_64ch leg): fused launches of up to 4 heads with the 4-block pass after "
           "them, tail chunks pinned to 8; the object's rows (100 MB) stay in the Infinity Cache"),
    "serving": ("k_group_tick_serving.txt",
                "the served mixed workload (Poisson sessions, 20/s, mean life 30 s, 1-16 channels, IR "
                "8,192-524,288 taps, B = 256, an IR update every 10-100 blocks; about 620 sessions / 5,300 "
                "channels), 4 s of audio after a 64-tick warm-up, uniform then non-uniform: one grouped IR "
                "update call and one group launch per kind per tick"),
    "room": ("room_responses_and_block_dft.txt",
             "the room-response generator (1 and 64 responses of 65,536 taps, both ears, per launch) and the "
             "block DFT on the matrix cores (1,024 frames of 256 and 512 samples, plain and split-bf16, beside "
             "the FP32 LDS FFT)"),
}


def profile_leg(T, torch, leg):
    """The workload of one profile summary other than the real-time rig's."""
    global TAIL_CHUNKS
    ctx = T.TANContext(device=0)
    if leg == "c2":
        TAIL_CHUNKS = 8  # the autotuner's choice in the unprofiled run (the tracer slows its dry runs)
        hs = torch.cuda.Stream()
        hctx = T.TANContext(device=0, conv_stream=hs, general_stream=hs)
        bench_throughput(T, torch, hctx, 2000, 100, timing=False, stream=hs)
        hctx.Terminate()
    elif leg == "serving":
        for meth in (UNIFORM_METHOD, NONUNIFORM_METHOD):
            bench_serving(T, torch, ctx, seconds=4.0, method=meth)
    elif leg == "room":
        bench_room(T, torch, ctx)
        bench_dft_mfma(T, torch, ctx, log2n=8)
        bench_dft_mfma(T, torch, ctx, log2n=9)
    ctx.Terminate()
    log(f"profile leg {leg} done")


def summarize_trace(run_dir, leg):
    """Writes profiles/<leg's file> from the kernel statistics of a traced
    `bench.py --profile-run --profile-leg <leg>` under run_dir/prof_<leg>."""
    import csv
    name, what = PROFILE_LEGS[leg]
    stats = list(csv.DictReader(open(os.path.join(run_dir, f"prof_{leg}", "run_kernel_stats.csv"))))
    cmd = (f"cd /tmp && export TMPDIR=/tmp && rocprofv3 --kernel-trace --stats --output-format csv -d "
           f"<repo>/gpurun_out/prof_{leg} -o run -- python3 <repo>/bench.py --profile-run --profile-leg {leg}")
    lines = [f"# Kernel trace: {what}", "", "Command (from the repository root on the GPU box):", "  " + cmd, "",
             f"{'kernel':64s} {'calls':>6s} {'avg us':>10s} {'min us':>10s} {'max us':>10s} {'% time':>7s}"]
    for r in stats:
        if float(r["Percentage"]) < 0.05:
            continue
        lines.append(f"{r['Name'].split('(')[0].replace('void ', '')[:64]:64s} {int(r['Calls']):6d} "
                     f"{float(r['AverageNs']) / 1e3:10.1f} {float(r['MinNs']) / 1e3:10.1f} "
                     f"{float(r['MaxNs']) / 1e3:10.1f} {float(r['Percentage']):7.2f}")
    out = os.path.join(ROOT, "profiles", name)
    with open(out, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return out


def is_group_kernel(name):
    """A tick's group launch in a rocprofv3 kernel name: the pass and tail
    kernels and the heads-only kernel."""
    return name.startswith("void tanx::k_group_tick") or name.startswith("void tanx::k_group_heads")


def pmc_traffic_per_channel():
    """HBM bytes per channel and tick of the real-time group launches (every
    k_group_tick and k_group_heads of a tick) from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE scaled by a 512 MiB copy), or None when it was
    taken with another pass length."""
    try:
        data = json.load(open(PMC_JSON))
    except (OSError, ValueError):
        return None
    if data.get("pass_blocks") != RT_PASS:
        return None
    return data.get("group_tick_traffic_bytes_per_channel")


def summarize_profiles(run_dir, groups):
    """Writes the committed summaries of the profiled rig runs (the commands
    in PROFILE_CMDS, outputs under run_dir): the kernel trace statistics as
    text and the PMC traffic as JSON, under profiles/."""
    import csv
    import collections
    channels = rt_channels(groups)
    pre = "cd /tmp && export TMPDIR=/tmp && "

    def counters(path, counter):
        per = collections.defaultdict(list)
        for r in csv.DictReader(open(path)):
            if r.get("Counter_Name") == counter:
                per[r["Kernel_Name"]].append((int(r["Grid_Size"]), float(r["Counter_Value"])))
        return per

    f = counters(os.path.join(run_dir, "pmc_fetch", "run_counter_collection.csv"), "FETCH_SIZE")
    w = counters(os.path.join(run_dir, "pmc_write", "run_counter_collection.csv"), "WRITE_SIZE")
    # the 512 MiB copy's counters (its large dispatches) calibrate KiB per byte
    cf = [v for g, v in f["__amd_rocclr_copyBuffer"] if g >= 131072]
    cw = [v for g, v in w["__amd_rocclr_copyBuffer"] if g >= 131072]
    kib = 512 * 1024
    ff, wf = kib / (sum(cf) / len(cf)), kib / (sum(cw) / len(cw))
    rows, tick_bytes = [], 0.0
    for name in sorted(f, key=lambda k: -sum(v for _, v in f[k])):
        if not name.startswith("void tanx::"):
            continue
        fv, wv = [v for _, v in f[name]], [v for _, v in w.get(name, [])]
        fk, wk = sum(fv) / len(fv), (sum(wv) / len(wv) if wv else 0.0)
        tb = fk * 1024 * ff + wk * 1024 * wf
        rows.append({"kernel": name, "dispatches": len(fv), "fetch_kib": round(fk, 1), "write_kib": round(wk, 1),
                     "traffic_bytes": round(tb)})
        if is_group_kernel(name):
            tick_bytes += tb * len(fv)
    per_tick = tick_bytes / PROFILE_TICKS
    pmc = {"source": f"rocprofv3 --pmc FETCH_SIZE, then --pmc WRITE_SIZE, of: python3 bench.py --profile-run "
                     f"--rt-groups {groups}",
           "commands": [pre + PROFILE_CMDS[k].format(g=groups) for k in ("fetch", "write")],
           "workload": f"real-time rig: {groups} groups of 2,048 channels ({channels} channels), 65,536 taps, "
                       f"B=256, {RT_PASS}-block passes, {PROFILE_TICKS} ticks with staggered passes, IR updates "
                       "and crossfades",
           "pass_blocks": RT_PASS, "channels": channels,
           "calibration": {"kernel": "__amd_rocclr_copyBuffer (torch copy_ of 512 MiB)", "bytes_read": 1 << 29,
                           "bytes_written": 1 << 29, "fetch_kib": sum(cf) / len(cf), "write_kib": sum(cw) / len(cw),
                           "fetch_factor": ff, "write_factor": wf},
           "kernels": rows,
           "group_tick_traffic_bytes_per_tick": round(per_tick),
           "group_tick_traffic_bytes_per_channel": per_tick / channels}
    with open(PMC_JSON, "w") as fh:
        json.dump(pmc, fh, indent=1)
    stats = list(csv.DictReader(open(os.path.join(run_dir, "prof_rt", "run_kernel_stats.csv"))))
    gt_ns = sum(float(r["TotalDurationNs"]) for r in stats if is_group_kernel(r["Name"]))
    lines = [f"# Kernel trace: real-time rig, {groups} groups x 2,048 channels ({channels:,} channels), "
             f"65,536 taps, B = 256, {RT_PASS}-block passes", "",
             "Command (from the repository root on the GPU box):",
             "  " + pre + PROFILE_CMDS["trace"].format(g=groups), "",
             f"Workload: {PROFILE_TICKS} ticks of the benchmark's real-time rig (one group launch per tick and role "
             f"set, one block in flight, pass phases g % {RT_PASS}, each group's IR update every 50 ticks, "
             "staggered, with its crossfade); the first ticks include cold caches (the Max column).", "",
             f"{'kernel':64s} {'calls':>6s} {'avg us':>10s} {'min us':>10s} {'max us':>10s} {'% time':>7s}"]
    for r in stats:
        lines.append(f"{r['Name'].split('(')[0].replace('void ', '')[:64]:64s} {int(r['Calls']):6d} "
                     f"{float(r['AverageNs']) / 1e3:10.1f} {float(r['MinNs']) / 1e3:10.1f} "
                     f"{float(r['MaxNs']) / 1e3:10.1f} {float(r['Percentage']):7.2f}")
    lines += ["", f"Group kernels per tick: {gt_ns / PROFILE_TICKS / 1e3:.1f} us (all k_group_tick and "
              f"k_group_heads launches of the {PROFILE_TICKS} ticks / {PROFILE_TICKS}).",
              f"HBM traffic of the group kernels ({os.path.relpath(PMC_JSON, ROOT)}): {per_tick / 1e9:.2f} GB per "
              f"tick, {per_tick / channels / 1e3:.1f} KB per channel (FETCH_SIZE x {ff:.3f} + WRITE_SIZE x "
              f"{wf:.3f}, the factors of a 512 MiB copy)."]
    txt = os.path.join(ROOT, "profiles", f"k_group_tick_realtime_{channels}ch_65536taps_b256.txt")
    with open(txt, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return txt


def main():
    global TAIL_CHUNKS, RT_PASS, RT_GROUP, RT_SLOTS
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed steps of 100 ticks (2,000 ticks at 20)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps of 100 ticks before them")
    ap.add_argument("--latency-blocks", type=int, default=300)
    ap.add_argument("--cpu-seconds", type=float, default=8.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tail-chunks", type=int, default=0,
                    help="pin the 64-channel passes' tail chunks instead of autotuning (for profiled runs)")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-batching", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--batching-streams", type=str, default="512,2048")
    ap.add_argument("--serving-seconds", type=float, default=15.0)
    ap.add_argument("--no-nonuniform", action="store_true")
    ap.add_argument("--no-room", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--serving-methods", type=str, default="uniform,nonuniform")
    ap.add_argument("--nu-sweep", type=str, default="65536,98304")
    ap.add_argument("--rt-groups", type=int, default=0,
                    help="profile run: real-time rig with this many groups of 2,048 channels")
    ap.add_argument("--rt-group-channels", type=int, default=RT_GROUP,
                    help="channels per object of the real-time count")
    ap.add_argument("--rt-pass-blocks", type=int, default=RT_PASS,
                    help="blocks per tail pass of the real-time groups (1..16)")
    ap.add_argument("--rt-ir-slots", type=int, default=RT_SLOTS, choices=(2, 3),
                    help="IR spectrum slots per real-time group")
    ap.add_argument("--summarize-profiles", type=str, default="",
                    help="write profiles/ summaries from the rocprofv3 outputs under this folder (no GPU)")
    ap.add_argument("--profile-leg", type=str, default="rt", choices=["rt"] + sorted(PROFILE_LEGS),
                    help="with --profile-run: the real-time rig (rt) or another leg's workload; with "
                         "--summarize-profiles: which summary to write")
    ap.add_argument("--serving-only", action="store_true",
                    help="run the serving legs alone and print their rows as one JSON line (no headline)")
    ap.add_argument("--profile-run", action="store_true",
                    help="short fixed workload for rocprofv3 counter passes; prints no JSON")
    args = ap.parse_args()
    TAIL_CHUNKS = args.tail_chunks
    RT_PASS = args.rt_pass_blocks
    RT_GROUP = args.rt_group_channels
    RT_SLOTS = args.rt_ir_slots
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (channels and "
              "sessions shard across GPUs without exchange; see the design notes)", file=sys.stderr)
        sys.exit(2)

    if args.summarize_profiles:
        if args.profile_leg == "rt":
            print(summarize_profiles(args.summarize_profiles, args.rt_groups or PROFILE_GROUPS))
        else:
            print(summarize_trace(args.summarize_profiles, args.profile_leg))
        return
    import torch
    import tan_amd as T
    T.load_library()
    if args.profile_run:
        if args.profile_leg == "rt":
            profile_run(T, torch, args.rt_groups or PROFILE_GROUPS)
        else:
            profile_leg(T, torch, args.profile_leg)
        return
    if args.serving_only:
        sctx = T.TANContext(device=0)
        names = {"uniform": UNIFORM_METHOD, "nonuniform": NONUNIFORM_METHOD}
        rows = []
        for meth in [names[m] for m in args.serving_methods.split(",") if m]:
            rows.append(bench_serving(T, torch, sctx, seconds=args.serving_seconds, method=meth))
            log(f"  {rows[-1]}")
        sctx.Terminate()
        print(json.dumps({"serving": rows}))
        return

    log("real-time channel count (value)")
    rt = bench_realtime(T, torch, args.steps, args.warmup)
    value = rt["channels"]
    gt = rt["group_timing"]
    launches = max(1, gt["tick_launches"])
    tick_us = gt["tick_ms"] / launches * 1e3
    achieved = gt["tick_alg_bytes"] / (gt["tick_ms"] * 1e-3) / 1e9 if gt["tick_ms"] > 0 else 0.0
    per_ch = pmc_traffic_per_channel()
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBPS,
                "traffic": per_ch * value if per_ch else None,
                "kernel": f"one tick's group launches: k_group_heads<0, 256> (the heads of the groups "
                          f"without a pass this tick, one per wave), k_group_tick<0, 256, 1, {RT_PASS}> (the groups whose "
                          f"single-filter pass runs this tick, one in {RT_PASS}, and the new filter's catch-up "
                          f"pass of each group whose IR switched, with their heads), k_group_tick<0, 256, 2, "
                          f"{RT_PASS}> when a switch falls on a group's pass block (its two-filter pass)",
                "trace": f"profiles/k_group_tick_realtime_{rt_channels(PROFILE_GROUPS)}ch_65536taps_b256.txt",
                "traffic_source": "profiles/pmc_traffic_k_group_tick_rt.json" if per_ch else None,
                "alg_bytes_per_launch": gt["tick_alg_bytes"] / launches, "avg_launch_us": tick_us,
                "launches": gt["tick_launches"], "channels": value}
    log(f"  {value} real-time channels; group kernel {tick_us:.1f} us per launch, {achieved:.0f} GB/s algorithmic")

    ctx = T.TANContext(device=0)
    # The 64-channel pass uses a context created on a torch stream (the
    # caller-stream form of the reference's InitOpenCL(queue, queue)): the
    # inputs are written on that stream, so Process needs no cross-stream
    # wait per call, as a C++ caller on its own stream would not.
    hs = torch.cuda.Stream()
    hctx = T.TANContext(device=0, conv_stream=hs, general_stream=hs)

    c2 = None
    if not args.no_c2:
        log("C2 workload (64 channels, 2,000 blocks, updates every 50)")
        dt, _, info, finite = bench_throughput(T, torch, hctx, 2000, 100, timing=False, stream=hs)
        assert finite, "non-finite output"
        dt2, kt, info2, _ = bench_throughput(T, torch, hctx, 2000, 100, timing=True, stream=hs)
        pass_us = kt["tail_ms"] / max(1, kt["tail_launches"]) * 1e3
        c2 = {"channels": C, "blocks": 2000, "ir_updates": 40, "wall_ms": dt * 1e3,
              "rt_channels_throughput": C * 2000 * B / FS / dt, "ms_per_update_period": dt / 40 * 1e3,
              "pass_blocks": info.get("pass_blocks"), "tail_chunks": info.get("tail_chunks"),
              "k_window_avg_us": pass_us,
              "k_window_alg_GBps": kt["tail_alg_bytes"] / (kt["tail_ms"] * 1e-3) / 1e9 if kt["tail_ms"] else None,
              "kernel_us_per_block": (kt["head_ms"] + kt["tail_ms"]) * 1e3 / kt["blocks"],
              "note": "64-channel rows (~100 MB) stay in the Infinity Cache; not an HBM figure"}
        log(f"  {c2['rt_channels_throughput']:.0f} real-time channels of throughput at 64 channels")

    log("latency pass")
    lat = bench_latency(T, ctx, args.latency_blocks)
    log(f"  {lat}")

    prec_rows, dft_rows = [], []
    if not args.no_precision:
        log("precision pass")
        prec_rows = bench_precision(T, torch, ctx)
        # 256 points (the configuration's block DFT) and 512 (the head's 2B-point frame)
        dft_rows = bench_dft_mfma(T, torch, ctx, log2n=8) + bench_dft_mfma(T, torch, ctx, log2n=9)

    nonuniform = None
    if not args.no_nonuniform:
        log("non-uniform pass")
        dtn, _, infon, finite_n = bench_throughput(T, torch, hctx, 2000, 100, timing=False,
                                                   method=NONUNIFORM_METHOD, stream=hs)
        assert finite_n, "non-finite output (non-uniform)"
        nu_value = C * 2000 * B / FS / dtn
        log(f"  non-uniform (M = {infon['nu_multiple']}): {nu_value:.0f} real-time channels of throughput")
        counts = [int(v) for v in args.nu_sweep.split(",") if v]
        ph_best, ph_rows = bench_sweep_phased(T, torch, ctx, counts)
        nonuniform = {"method": "TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM",
                      "nu_multiple": infon["nu_multiple"], "rt_channels_throughput_64ch": nu_value,
                      "rt_channels_at_deadline_phased_no_updates": ph_best, "sweep_phased": ph_rows}

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

    out = {
        "metric": METRIC, "value": value, "unit": "real-time channels", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": rt["wall_s"] / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (decaying-noise IRs, U[-1,1] input, generated on device)",
        "config": {"workload": RT_WORKLOAD, "channels": value, "channel_groups": rt["groups"],
                   "group_channels": RT_GROUP, "large_groups": min(rt["groups"], rt_large()),
                   "small_group_channels": RT_GROUP // 2, "block": B, "ir_taps": L, "ticks": rt["ticks"],
                   "ticks_per_step": RT_STEP_TICKS, "update_every": UPDATE_EVERY, "pass_blocks": RT_PASS,
                   "ir_slots": RT_SLOTS, "tail_chunks": rt_chunks(),
                   "sample_rate": FS, "deadline_ms": DEADLINE_MS},
        "deadline_misses": rt["misses"],
        "p50_block_ms": rt["p50_host_ms"], "p99_block_ms": rt["p99_host_ms"], "max_block_ms": rt["max_host_ms"],
        "p50_device_ms": rt["p50_device_ms"], "p99_device_ms": rt["p99_device_ms"],
        "max_device_ms": rt["max_device_ms"],
        "rt_throughput_channels": value * rt["ticks"] * B / FS / rt["wall_s"],
        "roofline": roofline,
        "realtime_search": rt["probes"], "realtime_outputs_finite": rt["outputs_finite"],
        "c2_64ch": c2,
        "latency_64ch_host_buffers": lat,
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
