"""Batched streams of uneven IR length and block size in one launch per tick.

Every stream is its own TANConvolution object (its own FDL, overlap and IR
slots in HBM); the streams that are due at a base tick advance one block
each in ONE group kernel launch (tanx_group_process_device).  The base tick
is the smallest block size; a stream of block B runs every B / base ticks,
and streams of one block size are spread evenly over the phases so that each
base tick carries the same share of the long-block work.

The reference runs one engine per TANConvolution object
(ConvolutionImpl.cpp:1413-1767) and has no batching across objects; the
streams' results here are the same as each object processed alone, because
objects keep independent state (checked by tests/test_gpu_group.py and
tests/test_gpu_batching.py).

The tuner picks the tail chunk count (partitions per workgroup of the
partial-sum multiply-accumulate) per (block size, rounded IR length) bucket
by coordinate descent on the measured device time of whole tick cycles.
"""
import ctypes
import math

import numpy as np

from . import tan as T

SAMPLE_RATE = 48000


class GroupCall:
    """Argument arrays of tanx_group_process_device, built once.

    convs: TANConvolution objects of one context; inputs/outputs: torch CUDA
    tensors (channels, >= block) per object."""

    def __init__(self, convs, inputs, outputs):
        n = len(convs)
        for c, i, o in zip(convs, inputs, outputs):
            # one block per row is read and written through pointer and row stride
            for t in (i, o):
                if t.dim() != 2 or t.shape[0] < c.channels or t.shape[1] < c.block:
                    raise ValueError("group input/output views must hold (channels, >= block) samples")
        self.n = n
        self._lib = convs[0]._lib if n else None
        self.h = (ctypes.c_void_p * n)(*[c.handle for c in convs])
        self.i = (ctypes.c_void_p * n)(*[t.data_ptr() for t in inputs])
        self.o = (ctypes.c_void_p * n)(*[t.data_ptr() for t in outputs])
        self.ist = (ctypes.c_size_t * n)(*[t.stride(0) for t in inputs])
        self.ost = (ctypes.c_size_t * n)(*[t.stride(0) for t in outputs])
        # keep the tensors alive as long as the pointers are in use
        self._keep = (list(convs), list(inputs), list(outputs))
        self._ctx = convs[0]._ctx if n else None

    def launch(self, order=True):
        """One group launch; order: first make the process stream wait for
        torch's current stream (False when the caller has ordered it)."""
        if self.n == 0:
            return T.AMF_OK
        if order:
            self._ctx.order_after_torch()
        return self._lib.tanx_group_process_device(self.n, self.h, self.i, self.ist, self.o, self.ost, None)


def rounded_length(ir_len, block):
    """The object's internal IR length: next power of two >= ir_len and >= block
    (ConvolutionImpl.cpp:1137-1143)."""
    n = 1
    while n < ir_len:
        n <<= 1
    return max(n, block)


def alg_bytes_per_block(ir_len, block, channels=1, multiple=1):
    """Algorithmic HBM bytes of one block (FP32 spectra).

    Uniform (multiple = 1): IR + FDL reads of all partitions, the new FDL
    row, input, output and overlap history: 16*P*K + 8*K + 16*B per channel,
    K = B + 1 (the reference's Hermitian bins).
    Non-uniform (frames of N = multiple * B): the first frame of taps as
    uniform partitions (P = multiple), one slice of B bins of frame partitions
    2..PB-1 per block (16*(PB-2)*B), and the frame kernel's input frame, frame
    spectrum written and previous one read, two frame-partition spectra and
    the frame output, spread over the frame's blocks."""
    k = block + 1
    if multiple <= 1:
        p = rounded_length(ir_len, block) // block
        return channels * (16 * p * k + 8 * k + 16 * block)
    n = multiple * block
    pb = rounded_length(ir_len, block) // n - 1
    head = 16 * multiple * k + 8 * k + 16 * block
    mac = 16 * (pb - 2) * block if pb > 2 else 0
    frame = (2 * n * 4 + 2 * n * 8 + 2 * n * 8 + n * 4) // multiple
    return channels * (head + mac + frame)


def draw_streams(n, seed=7, min_len=4096, max_len=262144, blocks=(128, 256, 512)):
    """n (ir_len, block) pairs: lengths log-uniform in [min_len, max_len],
    blocks uniform over `blocks` (a seeded PCG64 stream)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    lens = np.exp(rng.uniform(math.log(min_len), math.log(max_len), n)).astype(np.int64)
    bl = rng.choice(np.asarray(blocks), n)
    return [(int(a), int(b)) for a, b in zip(lens, bl)]


def phase_plan(blocks):
    """Base block, period (in base ticks) and phase of every stream.

    Streams of one block size take phases 0, 1, ..., period-1 in turn, so each
    base tick carries 1/period of them."""
    base = min(blocks)
    periods = [b // base for b in blocks]
    seen = {}
    phases = []
    for per in periods:
        k = seen.get(per, 0)
        phases.append(k % per)
        seen[per] = k + 1
    cycle = max(periods)
    return base, periods, phases, cycle


NONUNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM
UNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED


def layout_candidates(block, ir_len):
    """Partition layouts the tuner tries for streams of this block size and IR
    length: uniform partitions of B with 4- or 8-block passes, and frames of
    M x B for the non-uniform method (M = 2..16, a frame of at most 4,096
    samples and at most half the IR), as the reference's bestNUMultiple
    (ConvolutionImpl.cpp:224-246) chooses among multiples, here by
    measurement instead of a cost model."""
    ln = rounded_length(ir_len, block)
    out = [("uniform", 4), ("uniform", 8)] if ln // block >= 16 else [("uniform", 1)]
    m = 2
    while m <= 16 and m * block <= 4096 and 2 * m * block <= ln:
        out.append(("nonuniform", m))
        m *= 2
    return out


class StreamBatch:
    """Streams of uneven (ir_len, block) advancing together on one device.

    specs: list of (ir_len, block[, channels]) per stream."""

    def __init__(self, ctx, specs, ir_precision=T.TAN_IR_PRECISION_FP32, device="cuda", ir_seed=11,
                 method=T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED):
        import torch

        self.ctx = ctx
        self.specs = [(int(s[0]), int(s[1]), int(s[2]) if len(s) > 2 else 1) for s in specs]
        blocks = [b for _, b, _ in self.specs]
        for b in blocks:
            if b & (b - 1) or b < 64 or b > 512:
                raise ValueError("stream block sizes must be powers of two in [64, 512]")
        self.base, self.periods, self.phases, self.cycle = phase_plan(blocks)
        if any(self.cycle % p for p in self.periods):
            raise ValueError("block sizes must divide the largest block")
        self.convs = []
        self.inputs = []
        self.outputs = []
        self.irs = []
        self.layouts = []  # per stream: (kind, pass length or frame multiple); None = the library's choice
        self.buckets = {}
        self.ir_precision = ir_precision
        self.method = method
        gen = torch.Generator(device=device)
        gen.manual_seed(ir_seed)
        for j, (ln, b, c) in enumerate(self.specs):
            conv = T.TANConvolution(ctx, ir_precision)
            rc = conv.InitGpu(method, ln, b, c)
            if rc != T.AMF_OK:
                raise RuntimeError(f"InitGpu(L={ln}, B={b}, C={c}) failed: AMF_RESULT {rc}")
            # decaying-noise IR (exp(-t / 0.25 s)), unit L2, generated on the device
            t = torch.arange(ln, device=device, dtype=torch.float32)
            h = torch.randn(c, ln, device=device, generator=gen) * torch.exp(-t / (0.25 * SAMPLE_RATE))
            h = h / h.norm(dim=1, keepdim=True)
            # the IR transform reads h on the library's update stream
            torch.cuda.current_stream().synchronize()
            rc = conv.UpdateResponseTD(h, ln, None, T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY)
            if rc != T.AMF_OK:
                raise RuntimeError(f"UpdateResponseTD failed: AMF_RESULT {rc}")
            self.convs.append(conv)
            self.irs.append(h)
            self.layouts.append(None)
            self.inputs.append(torch.empty(c, b, device=device).uniform_(-1.0, 1.0, generator=gen))
            self.outputs.append(torch.zeros(c, b, device=device))
            self.buckets.setdefault((b, rounded_length(ln, b)), []).append(j)
        self.due = [[j for j in range(len(self.specs)) if t % self.periods[j] == self.phases[j]]
                    for t in range(self.cycle)]
        self._build_calls()
        self.bytes_per_tick = [sum(alg_bytes_per_block(*self.specs[j]) for j in d) for d in self.due]
        self.ticks = 0
        self._stream = None

    def _build_calls(self):
        self.calls = [GroupCall([self.convs[j] for j in d], [self.inputs[j] for j in d],
                                [self.outputs[j] for j in d]) for d in self.due]

    def stream_options(self, j):
        """(method, TANConvolution keyword arguments) that rebuild stream j's
        object as the batch has it."""
        lay = self.layouts[j]
        if lay is None:
            return self.method, {"ir_precision": self.ir_precision}
        kind, v = lay
        if kind == "uniform":
            # the bucket's streams take pass phases 0..v-1 in turn, so a
            # bucket's passes are spread over its blocks
            phase = self.buckets[(self.specs[j][1], rounded_length(self.specs[j][0], self.specs[j][1]))].index(j) % v
            return UNIFORM, {"ir_precision": self.ir_precision, "pass_blocks": v, "pass_phase": phase}
        return NONUNIFORM, {"ir_precision": self.ir_precision, "nu_multiple": v}

    def set_bucket_layout(self, key, layout):
        """Rebuilds the objects of one (block, rounded length) bucket with a
        partition layout of layout_candidates (their state starts over: the
        IR is transformed again and fades in)."""
        import torch
        for j in self.buckets[key]:
            ln, b, c = self.specs[j]
            self.layouts[j] = layout
            method, kw = self.stream_options(j)
            conv = T.TANConvolution(self.ctx, **kw)
            rc = conv.InitGpu(method, ln, b, c)
            if rc != T.AMF_OK:
                raise RuntimeError(f"InitGpu(L={ln}, B={b}, C={c}, {layout}) failed: AMF_RESULT {rc}")
            torch.cuda.current_stream().synchronize()
            rc = conv.UpdateResponseTD(self.irs[j], ln, None, T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY)
            if rc != T.AMF_OK:
                raise RuntimeError(f"UpdateResponseTD failed: AMF_RESULT {rc}")
            self.convs[j].Terminate()
            self.convs[j] = conv
        self._build_calls()

    def reset(self):
        """Rebuilds every object with its current layout: every stream starts
        over from block 0 (after tuning, whose timed ticks advanced them)."""
        for key, js in self.buckets.items():
            self.set_bucket_layout(key, self.layouts[js[0]])

    @property
    def base_period_ms(self):
        return 1e3 * self.base / SAMPLE_RATE

    def stream(self):
        import torch
        if self._stream is None:
            self._stream = torch.cuda.ExternalStream(self.convs[0].ProcessStream())
        return self._stream

    def tick(self, after_current=True):
        """Enqueue one base tick: every due stream advances one block.

        With after_current, the launch waits for work already queued on
        torch's current stream (e.g. copies into the input blocks)."""
        if after_current:
            import torch
            ev = torch.cuda.Event()
            ev.record()
            self.stream().wait_event(ev)
        rc = self.calls[self.ticks % self.cycle].launch()
        if rc != T.AMF_OK:
            raise RuntimeError(f"group launch failed: AMF_RESULT {rc}")
        self.ticks += 1

    def time_ticks(self, n, per_tick=False):
        """Device time of n consecutive base ticks, in ms (total, or a list
        per tick with a synchronisation after each)."""
        import torch
        s = self.stream()
        if per_tick:
            out = []
            for _ in range(n):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(s)
                self.tick(after_current=False)
                e1.record(s)
                e1.synchronize()
                out.append(e0.elapsed_time(e1))
            return out
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(n):
            self.tick(after_current=False)
        e1.record(s)
        e1.synchronize()
        return e0.elapsed_time(e1)

    def set_bucket_chunks(self, key, chunks):
        for j in self.buckets[key]:
            rc = self.convs[j].SetTailChunks(chunks)
            if rc != T.AMF_OK:
                raise RuntimeError(f"SetTailChunks failed: AMF_RESULT {rc}")

    def _best_of(self, n, reps=3):
        """Least device time of `reps` runs of n ticks (ms)."""
        return min(self.time_ticks(n) for _ in range(reps))

    def tune_layouts(self, cycles=4, min_gain=0.03):
        """Coordinate descent over buckets (largest byte share first): each
        bucket measures its current layout, then every partition layout of
        layout_candidates with the others fixed, and switches only to one
        whose device time over `cycles` full tick cycles is at least
        `min_gain` lower (a tick near its host enqueue time measures with
        noise of a few percent).  Returns {bucket: layout} (None: the
        library's choice)."""
        n = cycles * self.cycle
        share = {k: sum(alg_bytes_per_block(*self.specs[j]) // self.periods[j] for j in v)
                 for k, v in self.buckets.items()}
        chosen = {}
        self.time_ticks(n)  # warm-up
        for k in sorted(self.buckets, key=lambda k: -share[k]):
            cur = self.layouts[self.buckets[k][0]]
            best = (self._best_of(n), cur)
            for lay in layout_candidates(k[0], k[1]):
                if lay == cur:
                    continue
                self.set_bucket_layout(k, lay)
                self.time_ticks(self.cycle)  # past the rebuilt objects' first pass
                ms = self._best_of(n)
                if ms < best[0] * (1.0 - min_gain):
                    best = (ms, lay)
            chosen[k] = best[1]
            if self.layouts[self.buckets[k][0]] != best[1]:
                self.set_bucket_layout(k, best[1])
        return chosen

    def tune(self, cycles=4, candidates=(1, 2, 4, 8, 16, 32, 64), passes=1, min_gain=0.03):
        """Coordinate descent over buckets (largest byte share first): each
        bucket measures its current chunk count, then every candidate with
        the others fixed, and switches only to one whose device time over
        `cycles` full tick cycles is at least `min_gain` lower.  Returns
        {bucket: chunks} (0: the library's choice)."""
        n = cycles * self.cycle
        share = {k: sum(alg_bytes_per_block(*self.specs[j]) // self.periods[j] for j in v)
                 for k, v in self.buckets.items()}
        order = sorted(self.buckets, key=lambda k: -share[k])
        chosen = {k: 0 for k in order}  # library default
        self.time_ticks(n)  # warm-up
        for _ in range(passes):
            for k in order:
                b, ln = k
                parts = max(1, ln // b - 2)
                best = (self._best_of(n), chosen[k])
                for c in candidates:
                    if c > parts or c == chosen[k]:
                        continue
                    self.set_bucket_chunks(k, c)
                    ms = self._best_of(n)
                    if ms < best[0] * (1.0 - min_gain):
                        best = (ms, c)
                chosen[k] = best[1]
                self.set_bucket_chunks(k, best[1])
        return chosen

    def restore_defaults(self):
        """Every bucket back to the library's layout and chunk choice (the
        objects are rebuilt and start over)."""
        for key in self.buckets:
            self.set_bucket_layout(key, None)
