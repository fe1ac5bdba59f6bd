"""Continuous-batching convolution server on one device.

Sessions (a channel count, an IR length, IR updates over their lifetime)
arrive and leave at any time.  A session is admitted when its device bytes
fit the HBM budget; until then it waits in a FIFO queue, and a waiting
session that fits is admitted even when an older, larger one does not
(first fit in arrival order).  Admitted sessions keep their FDL, overlap and
IR slots resident in HBM (one TANConvolution object each); every audio tick,
all admitted sessions advance one block in ONE group kernel launch, while IR
transforms of new sessions and of IR updates run on the update stream and
overlap the ticks.

A session's output is the same as its object processed alone: objects keep
independent state and the group kernel runs the same per-object block code
(tests/test_gpu_serving.py checks this in a mix with arrivals, departures
and updates).

The reference has no server: one application drives each TANConvolution
object (ConvolutionImpl.cpp:848-1043, TALibTestDynamicChannelConvolution.cpp:
159-241 drives channel churn inside one object).  Block latency here is
measured from the start of a tick's host work to the completion of its
kernel on the device, with events on the process stream.
"""
import collections
import ctypes
import math
import time

import numpy as np

from . import _lib
from . import tan as T
from .batching import GroupCall, alg_bytes_per_block

SAMPLE_RATE = 48000


def estimate_bytes(lib, ir_len, block, channels, ir_precision=T.TAN_IR_PRECISION_FP32,
                   method=T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED, nu_multiple=0):
    """HBM bytes of one session: the object (tanx_conv_estimate_bytes) plus its
    input and output block."""
    opt = _lib.ConvOptions()
    opt.ir_precision = ir_precision
    opt.nu_multiple = nu_multiple
    out = ctypes.c_uint64(0)
    rc = lib.tanx_conv_estimate_bytes(method, ir_len, block, channels, ctypes.addressof(opt), ctypes.byref(out))
    if rc != T.AMF_OK:
        raise ValueError(f"estimate failed for L={ir_len}, B={block}, C={channels}: AMF_RESULT {rc}")
    return int(out.value) + 2 * channels * block * 4


class Admission:
    """HBM-budget admission: FIFO queue, first fit in arrival order.

    Pure bookkeeping (no device work), so that the policy is testable on its
    own: `submit` queues a request of `nbytes`, `release` returns bytes, and
    `admit` pops every queued request that fits now."""

    def __init__(self, budget_bytes):
        self.budget = int(budget_bytes)
        self.used = 0
        self.queue = collections.deque()
        self.admitted = 0
        self.rejected = 0

    def submit(self, key, nbytes):
        if nbytes > self.budget:
            self.rejected += 1
            return False
        self.queue.append((key, int(nbytes)))
        return True

    def admit(self):
        out = []
        if not self.queue:
            return out
        keep = collections.deque()
        for key, nb in self.queue:
            if self.used + nb <= self.budget:
                self.used += nb
                self.admitted += 1
                out.append((key, nb))
            else:
                keep.append((key, nb))
        self.queue = keep
        return out

    def release(self, nbytes):
        self.used -= int(nbytes)
        if self.used < 0:
            raise RuntimeError("released more bytes than admitted")

    def waiting(self):
        return len(self.queue)


class RowAllocator:
    """Contiguous row ranges of an I/O arena: first fit over address-ordered
    free intervals, coalesced on free.  A session's channels must be equally
    spaced rows (the group launch takes one stride per object)."""

    def __init__(self, rows):
        self.rows = int(rows)
        self.free = [(0, self.rows)]  # (start, length), sorted, disjoint, non-adjacent

    def alloc(self, n):
        for i, (st, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    del self.free[i]
                else:
                    self.free[i] = (st + n, ln - n)
                return st
        return None

    def release(self, start, n):
        import bisect
        i = bisect.bisect_left(self.free, (start, 0))
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] > start:
            raise RuntimeError("double free of arena rows")
        if i < len(self.free) and start + n > self.free[i][0]:
            raise RuntimeError("double free of arena rows")
        st, ln = start, n
        if i < len(self.free) and st + ln == self.free[i][0]:
            ln += self.free[i][1]
            del self.free[i]
        if i > 0 and self.free[i - 1][0] + self.free[i - 1][1] == st:
            st = self.free[i - 1][0]
            ln += self.free[i - 1][1]
            del self.free[i - 1]
            i -= 1
        self.free.insert(i, (st, ln))

    def free_rows(self):
        return sum(ln for _, ln in self.free)


class Session:
    __slots__ = ("sid", "channels", "ir_len", "conv", "inp", "out", "row", "nbytes", "blocks", "arrived_tick",
                 "admitted_tick", "updates", "multiple")

    def __init__(self, sid, channels, ir_len):
        self.sid = sid
        self.channels = channels
        self.ir_len = ir_len
        self.conv = None
        self.inp = None
        self.out = None
        self.row = None
        self.multiple = 1
        self.nbytes = 0
        self.blocks = 0
        self.arrived_tick = 0
        self.admitted_tick = -1
        self.updates = 0


class TickRecord:
    __slots__ = ("tick", "sessions", "channels", "latency_ms", "enqueue_ms", "alg_bytes")

    def __init__(self, tick, sessions, channels, latency_ms, enqueue_ms, alg_bytes):
        self.tick = tick
        self.sessions = sessions
        self.channels = channels
        self.latency_ms = latency_ms
        self.enqueue_ms = enqueue_ms
        self.alg_bytes = alg_bytes


class SessionServer:
    """Sessions of one block size on one context, one group launch per tick.

    The caller writes each admitted session's input block into `session.inp`
    (a (channels, block) CUDA tensor) before `tick()`, and reads
    `session.out` after it.  With io_rows, those are row ranges of two
    (io_rows, block) arenas (`in_arena`, `out_arena`), so that one copy per
    tick can bring in every session's input; a session that finds no free
    range gets tensors of its own."""

    def __init__(self, ctx, block=256, hbm_budget=32 << 30, ir_precision=T.TAN_IR_PRECISION_FP32,
                 device="cuda", io_rows=0, method=T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED,
                 nu_multiple=0):
        if block & (block - 1) or block < 64 or block > 512:
            raise ValueError("block must be a power of two in [64, 512]")
        self.ctx = ctx
        self.lib = ctx._lib
        self.block = block
        self.ir_precision = ir_precision
        self.method = method
        self.nu_multiple = nu_multiple
        self.device = device
        self.admission = Admission(hbm_budget)
        self.pending = {}
        self.active = collections.OrderedDict()
        self.ticks = 0
        self._call = None
        self._bytes_tick = 0
        self._chan_tick = 0
        self._stream = None
        self.rows = RowAllocator(io_rows) if io_rows else None
        self.in_arena = self.out_arena = None
        if io_rows:
            import torch
            self.in_arena = torch.zeros(io_rows, block, device=device)
            self.out_arena = torch.zeros(io_rows, block, device=device)

    @property
    def period_ms(self):
        return 1e3 * self.block / SAMPLE_RATE

    # -- control plane --
    def submit(self, sid, channels, ir_len, first_ir):
        """Queue a session; first_ir is a (channels, >= ir_len) CUDA tensor
        (kept until admission).  Returns False if it can never fit."""
        s = Session(sid, channels, ir_len)
        s.arrived_tick = self.ticks
        nb = estimate_bytes(self.lib, ir_len, self.block, channels, self.ir_precision, self.method,
                            self.nu_multiple)
        if not self.admission.submit(sid, nb):
            return False
        self.pending[sid] = (s, first_ir)
        return True

    def admit(self):
        """Create objects for every queued session that fits the budget now.
        Returns the admitted sessions."""
        import torch
        out = []
        for sid, nb in self.admission.admit():
            s, ir = self.pending.pop(sid)
            conv = T.TANConvolution(self.ctx, self.ir_precision, self.nu_multiple)
            rc = conv.InitGpu(self.method, s.ir_len, self.block, s.channels)
            if rc != T.AMF_OK:
                self.admission.release(nb)
                raise RuntimeError(f"InitGpu failed for session {sid}: AMF_RESULT {rc}")
            # the IR transform runs on the update stream; the session's first
            # blocks are silent and the IR fades in, as for any first
            # UpdateResponseTD (ConvolutionImpl.cpp:115-117, 877-991)
            rc = conv.UpdateResponseTD(ir, s.ir_len, None, T.TAN_CONVOLUTION_OPERATION_FLAG_NONE)
            if rc != T.AMF_OK:
                conv.Terminate()
                self.admission.release(nb)
                raise RuntimeError(f"UpdateResponseTD failed for session {sid}: AMF_RESULT {rc}")
            s.conv = conv
            s.multiple = conv.Info()["nu_multiple"]
            s.nbytes = nb
            r = self.rows.alloc(s.channels) if self.rows else None
            if r is not None:
                s.row = r
                s.inp = self.in_arena[r:r + s.channels]
                s.out = self.out_arena[r:r + s.channels]
            else:
                s.inp = torch.zeros(s.channels, self.block, device=self.device)
                s.out = torch.zeros(s.channels, self.block, device=self.device)
            s.admitted_tick = self.ticks
            self.active[sid] = s
            out.append(s)
        if out:
            self._call = None
        return out

    def update(self, sid, ir):
        """UpdateResponseTD for an admitted session (asynchronous; the
        crossfade happens two blocks later)."""
        s = self.active[sid]
        rc = s.conv.UpdateResponseTD(ir, s.ir_len, None, T.TAN_CONVOLUTION_OPERATION_FLAG_NONE)
        if rc != T.AMF_OK:
            raise RuntimeError(f"UpdateResponseTD failed for session {sid}: AMF_RESULT {rc}")
        s.updates += 1

    def release(self, sid):
        """End a session (queued or admitted) and return its bytes."""
        if sid in self.pending:
            s, _ = self.pending.pop(sid)
            self.admission.queue = collections.deque((k, nb) for k, nb in self.admission.queue if k != sid)
            return s
        s = self.active.pop(sid)
        s.conv.Terminate()
        if s.row is not None:
            self.rows.release(s.row, s.channels)
            s.row = None
        self.admission.release(s.nbytes)
        self._call = None
        return s

    # -- data plane --
    def _rebuild(self):
        ss = list(self.active.values())
        self._call = GroupCall([s.conv for s in ss], [s.inp for s in ss], [s.out for s in ss])
        self._bytes_tick = sum(alg_bytes_per_block(s.ir_len, self.block, s.channels, s.multiple) for s in ss)
        self._chan_tick = sum(s.channels for s in ss)

    def stream(self):
        import torch
        if self._stream is None:
            h = self.lib.tanx_ctx_process_stream(self.ctx.handle)
            self._stream = torch.cuda.ExternalStream(h)
        return self._stream

    def tick(self, timed=True, after_current=True):
        """Advance every admitted session one block in one launch.

        With timed=True, returns a TickRecord whose latency is the device
        time from the tick's start marker (recorded before any host work of
        the tick, on an idle process stream) to the end of its kernel, so it
        includes the host enqueue time."""
        import torch
        # membership changes rebuild the launch arguments before the tick
        # starts: control work, as a server thread prepares it off the audio path
        if self._call is None:
            self._rebuild()
        t0 = time.perf_counter()
        s = self.stream()
        if timed:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(s)
        if after_current:
            # inputs written on torch's current stream are complete first
            ev = torch.cuda.Event()
            ev.record()
            s.wait_event(ev)
        rc = self._call.launch(order=False)
        if rc != T.AMF_OK:
            raise RuntimeError(f"group launch failed: AMF_RESULT {rc}")
        t1 = time.perf_counter()
        for ss in self.active.values():
            ss.blocks += 1
        self.ticks += 1
        if not timed:
            return None
        e1.record(s)
        return (e0, e1, TickRecord(self.ticks - 1, len(self.active), self._chan_tick, 0.0,
                                   1e3 * (t1 - t0), self._bytes_tick))

    @staticmethod
    def finish(pending):
        """Wait for a timed tick and fill in its latency."""
        if pending is None:
            return None
        e0, e1, rec = pending
        e1.synchronize()
        rec.latency_ms = e0.elapsed_time(e1)
        return rec

    def shutdown(self):
        for sid in list(self.active):
            self.release(sid)
        for sid in list(self.pending):
            self.release(sid)


class PoissonWorkload:
    """Session arrivals and lifetimes of the served mixed workload.

    Poisson arrivals at `rate` per second, exponential lifetimes of mean
    `mean_life` seconds, channels uniform in [1, max_channels], IR length
    log-uniform in [min_len, max_len], an IR update every U{upd_lo..upd_hi}
    blocks.  Everything is drawn from one seeded PCG64 stream, in tick order."""

    def __init__(self, block=256, rate=20.0, mean_life=30.0, max_channels=16, min_len=8192, max_len=524288,
                 upd_lo=10, upd_hi=100, seed=5):
        self.block = block
        self.rate = rate
        self.mean_life = mean_life
        self.max_channels = max_channels
        self.min_len = min_len
        self.max_len = max_len
        self.upd_lo = upd_lo
        self.upd_hi = upd_hi
        self.rng = np.random.Generator(np.random.PCG64(seed))
        self.tick_s = block / SAMPLE_RATE
        self.next_sid = 0

    def new_session(self, tick):
        r = self.rng
        sid = self.next_sid
        self.next_sid += 1
        ch = int(r.integers(1, self.max_channels + 1))
        ln = int(math.exp(r.uniform(math.log(self.min_len), math.log(self.max_len))))
        life = max(1, int(round(r.exponential(self.mean_life) / self.tick_s)))
        upd = int(r.integers(self.upd_lo, self.upd_hi + 1))
        return {"sid": sid, "channels": ch, "ir_len": ln, "end_tick": tick + life, "update_every": upd}

    def arrivals(self, tick):
        """Sessions arriving during this tick."""
        return [self.new_session(tick) for _ in range(int(self.rng.poisson(self.rate * self.tick_s)))]

    def initial_population(self):
        """A stationary population at tick 0 (Poisson with mean rate * mean
        life; exponential lifetimes are memoryless, so the remaining lifetimes
        have the same law)."""
        n = int(self.rng.poisson(self.rate * self.mean_life))
        return [self.new_session(0) for _ in range(n)]
