"""Continuous-batching convolution server on one device.

Sessions (a channel count, an IR length, IR updates over their lifetime)
arrive and leave at any time.  A session is admitted when its device bytes
fit the HBM budget; until then it waits in a FIFO queue, and a waiting
session that fits is admitted even when an older, larger one does not
(first fit in arrival order).  Admitted sessions keep their FDL, overlap and
IR slots resident in HBM (one TANConvolution object each); every audio tick,
all admitted sessions advance one block in ONE group kernel launch.  IR
transforms of new sessions and of the tick's IR updates run on the process
stream ahead of the tick's launch, all of one tick in one launch per
partition size (update_many).  Uniform sessions get pass phases that spread
their multi-block passes evenly over the ticks.

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
                   method=T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED, nu_multiple=0, pass_blocks=0,
                   max_tail_chunks=0, ir_slots=0):
    """HBM bytes of one session: the object (tanx_conv_estimate_bytes, with
    the options it will be created with) plus its input and output block."""
    opt = _lib.ConvOptions()
    opt.ir_precision = ir_precision
    opt.nu_multiple = nu_multiple
    opt.pass_blocks = pass_blocks
    opt.max_tail_chunks = max_tail_chunks
    opt.ir_slots = ir_slots
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
                 "admitted_tick", "updates", "multiple", "pass_blocks", "pass_phase", "pass_bytes", "pass_slots",
                 "max_tail_chunks")

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
        self.pass_blocks = 0   # the object's pass length K (tanx_conv_info.pass_blocks)
        self.pass_phase = 0    # the pass phase it was created with
        self.pass_bytes = 0    # bytes one of its passes streams
        self.max_tail_chunks = 0
        self.pass_slots = ()   # tick residues (mod PASS_CYCLE) its passes land on

    def conv_options(self):
        """TANConvolution keyword arguments that rebuild this session's object
        (the same pass length and phase)."""
        return {"pass_blocks": self.pass_blocks if self.pass_slots else 0, "pass_phase": self.pass_phase,
                "max_tail_chunks": self.max_tail_chunks}


PASS_CYCLE = 8  # common period of pass lengths 1, 2, 4 and 8


def pass_residue(load, k):
    """Tick residue r (mod k) for the passes of a new session of pass length
    k: the one whose PASS_CYCLE // k residues mod PASS_CYCLE carry the least
    pass bytes at their busiest (ties: the lowest r)."""
    cost = [max(load[r + j * k] for j in range(PASS_CYCLE // k)) for r in range(k)]
    return min(range(k), key=lambda i: (cost[i], i))


def pass_bytes(ir_len, block, channels, ir_precision=T.TAN_IR_PRECISION_FP32):
    """Bytes one multi-block pass of a uniform object streams: the FDL row and
    the IR row of partitions 2..P-1 of every channel."""
    parts = max(0, -(-ir_len // block) - 2)
    ir_elem = 8 if ir_precision == T.TAN_IR_PRECISION_FP32 else 4
    return channels * parts * block * (8 + ir_elem)

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

    def __init__(self, ctx, block=256, hbm_budget=96 << 30, ir_precision=T.TAN_IR_PRECISION_FP32,
                 device="cuda", io_rows=0, method=T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED,
                 nu_multiple=0, pass_blocks=0, balance_passes=True, group_updates=True, max_tail_chunks=4):
        """pass_blocks: pass length K of uniform sessions (0: the library's
        choice); balance_passes: give each uniform session the pass phase
        that puts its passes on the tick residue (mod PASS_CYCLE) with the
        fewest pass bytes so far, so that no tick streams every session's
        rows (a session's output does not depend on its phase);
        group_updates: the IR updates issued between two ticks (first IRs
        of admitted sessions included) go to the device in one grouped call
        before the next tick (GroupUpdateResponseTD), else one call each;
        max_tail_chunks: partition chunks per channel of a session's pass
        (every session's pass shares one wide launch, so a few suffice; it
        also sizes the partial rows each session keeps)."""
        if block & (block - 1) or block < 64 or block > 512:
            raise ValueError("block must be a power of two in [64, 512]")
        if pass_blocks not in (0, 1, 2, 4, 8):
            raise ValueError("pass_blocks must be 0, 1, 2, 4 or 8")
        self.pass_blocks = pass_blocks
        self.balance_passes = balance_passes
        self.group_updates = group_updates
        self.max_tail_chunks = max_tail_chunks
        self._updates = collections.OrderedDict()  # sid -> (session, IR tensor), issued before the next tick
        self.pass_load = [0] * PASS_CYCLE
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
    def session_bytes(self, channels, ir_len):
        """HBM bytes a session of this server takes (the admission charge)."""
        uniform = self.method == T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
        return estimate_bytes(self.lib, ir_len, self.block, channels, self.ir_precision, self.method,
                              self.nu_multiple, self.pass_blocks if uniform else 0, self.max_tail_chunks)

    def submit(self, sid, channels, ir_len, first_ir):
        """Queue a session; first_ir is a (channels, >= ir_len) CUDA tensor
        (kept until admission).  Returns False if it can never fit."""
        s = Session(sid, channels, ir_len)
        s.arrived_tick = self.ticks
        nb = self.session_bytes(channels, ir_len)
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
            conv, phase = self._create(s)
            if conv is None:
                self.admission.release(nb)
                raise RuntimeError(f"InitGpu failed for session {sid}: AMF_RESULT {phase}")
            s.conv = conv
            info = conv.Info()
            s.multiple = info["nu_multiple"]
            s.nbytes = nb
            self._place_passes(s, info["pass_blocks"], phase)
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
            # the first IR (transformed on the process stream before the
            # session's first block, which is silent while the IR fades in,
            # as after any first UpdateResponseTD, ConvolutionImpl.cpp:115-117,
            # 877-991)
            self._queue_update(s, ir)
            out.append(s)
        if out:
            self._call = None
        return out

    def _create(self, s):
        """The session's object, with the pass phase that balances pass
        bytes over the ticks; returns (conv, phase) or (None, AMF_RESULT)."""
        uniform = self.method == T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
        parts = -(-s.ir_len // self.block)
        k = self.pass_blocks or (4 if parts >= 16 else 1)
        phase = 0
        if uniform and self.balance_passes and k > 1:
            # block b of the session runs at tick admitted + b; its passes
            # end on blocks with (b + phase) % k == 0, i.e. on ticks
            # t = admitted - phase (mod k)
            phase = (self.ticks - pass_residue(self.pass_load, k)) % k
        conv = T.TANConvolution(self.ctx, self.ir_precision, self.nu_multiple,
                                pass_blocks=self.pass_blocks if uniform else 0, pass_phase=phase,
                                max_tail_chunks=self.max_tail_chunks)
        rc = conv.InitGpu(self.method, s.ir_len, self.block, s.channels)
        if rc != T.AMF_OK:
            return None, rc
        return conv, phase

    def _place_passes(self, s, k, phase):
        s.pass_blocks, s.pass_phase = k, phase
        s.max_tail_chunks = self.max_tail_chunks
        if k <= 1 or PASS_CYCLE % k:
            return
        r = (self.ticks - phase) % k
        s.pass_bytes = pass_bytes(s.ir_len, self.block, s.channels, self.ir_precision)
        s.pass_slots = tuple(r + j * k for j in range(PASS_CYCLE // k))
        for i in s.pass_slots:
            self.pass_load[i] += s.pass_bytes

    def update(self, sid, ir):
        """UpdateResponseTD for an admitted session (asynchronous; issued
        before the next tick, the crossfade happens two blocks later)."""
        s = self.active[sid]
        self._queue_update(s, ir)
        s.updates += 1

    def _queue_update(self, s, ir):
        if not self.group_updates:
            self._issue([(s, ir)])
            return
        if s.sid in self._updates:  # a second update before the tick: keep the order
            self.flush_updates()
        self._updates[s.sid] = (s, ir)

    def _issue(self, items):
        convs = [s.conv for s, _ in items]
        irs = [ir[:s.channels, :s.ir_len] for s, ir in items]
        if len(items) == 1:
            rc = convs[0].UpdateResponseTD(irs[0], items[0][0].ir_len, None, T.TAN_CONVOLUTION_OPERATION_FLAG_NONE)
        else:
            rc = T.GroupUpdateResponseTD(convs, irs, [s.ir_len for s, _ in items])
        if rc != T.AMF_OK:
            raise RuntimeError(f"IR update failed for sessions {[s.sid for s, _ in items]}: AMF_RESULT {rc}")

    def flush_updates(self):
        """Issues the queued IR updates (one grouped call)."""
        if self._updates:
            items = list(self._updates.values())
            self._updates.clear()
            self._issue(items)

    def release(self, sid):
        """End a session (queued or admitted) and return its bytes."""
        if sid in self.pending:
            s, _ = self.pending.pop(sid)
            self.admission.queue = collections.deque((k, nb) for k, nb in self.admission.queue if k != sid)
            return s
        s = self.active.pop(sid)
        self._updates.pop(sid, None)
        for i in s.pass_slots:
            self.pass_load[i] -= s.pass_bytes
        s.pass_slots = ()
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
        # the queued IR updates run ahead of the tick on the process stream
        self.flush_updates()
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
