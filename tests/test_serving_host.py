"""Host logic of batching and serving, without a GPU: admission policy, the
workload generator's statistics, the phase plan of mixed block sizes, and the
byte model against its closed forms."""
import math

import numpy as np
import pytest

import tan_amd as T
from tan_amd import batching as Bt
from tan_amd import serving as S


def test_admission_first_fit_in_arrival_order():
    a = S.Admission(100)
    assert a.submit("big", 80) and a.submit("mid", 50) and a.submit("small", 20)
    got = a.admit()
    # "big" fits; "mid" does not beside it; "small" does and is not held back
    assert [k for k, _ in got] == ["big", "small"] and a.used == 100
    assert a.waiting() == 1
    assert a.admit() == []
    a.release(80)
    assert [k for k, _ in a.admit()] == ["mid"] and a.used == 70
    assert not a.submit("huge", 101) and a.rejected == 1
    with pytest.raises(RuntimeError):
        a.release(1000)


def test_admission_never_exceeds_budget():
    rng = np.random.default_rng(3)
    a = S.Admission(1000)
    live = {}
    for step in range(2000):
        if rng.random() < 0.6:
            a.submit(step, int(rng.integers(1, 400)))
        for k, nb in a.admit():
            live[k] = nb
        assert a.used == sum(live.values()) <= a.budget
        if live and rng.random() < 0.5:
            k = list(live)[int(rng.integers(len(live)))]
            a.release(live.pop(k))


def test_poisson_workload_statistics():
    w = S.PoissonWorkload(block=256, rate=20.0, mean_life=30.0, seed=9)
    ticks = 60000  # 320 s of audio
    n = sum(len(w.arrivals(t)) for t in range(ticks))
    expect = 20.0 * ticks * 256 / 48000
    assert abs(n - expect) < 5 * math.sqrt(expect)
    ss = [w.new_session(0) for _ in range(20000)]
    ch = np.array([s["channels"] for s in ss])
    assert ch.min() == 1 and ch.max() == 16 and abs(ch.mean() - 8.5) < 0.15
    ln = np.log(np.array([s["ir_len"] for s in ss], np.float64))
    lo, hi = math.log(8192), math.log(524288)
    assert ln.min() >= lo - 1e-9 and ln.max() <= hi
    assert abs(ln.mean() - (lo + hi) / 2) < 0.05  # log-uniform
    life_s = np.array([s["end_tick"] for s in ss]) * 256 / 48000
    assert abs(life_s.mean() - 30.0) < 1.0
    up = np.array([s["update_every"] for s in ss])
    assert up.min() == 10 and up.max() == 100
    pop = [len(S.PoissonWorkload(seed=s).initial_population()) for s in range(40)]
    assert abs(np.mean(pop) - 600) < 5 * math.sqrt(600 / 40)


def test_phase_plan_spreads_long_blocks():
    blocks = [128, 256, 512, 256, 512, 512, 512, 128, 256]
    base, periods, phases, cycle = Bt.phase_plan(blocks)
    assert base == 128 and cycle == 4
    assert periods == [1, 2, 4, 2, 4, 4, 4, 1, 2]
    for b, p, ph in zip(blocks, periods, phases):
        assert 0 <= ph < p
    # every stream is due exactly once per its period, and the 512-blocks
    # occupy distinct phases
    assert sorted(ph for b, ph in zip(blocks, phases) if b == 512) == [0, 1, 2, 3]
    for t in range(cycle):
        due = [j for j in range(len(blocks)) if t % periods[j] == phases[j]]
        assert all(blocks[j] == 128 for j in due if periods[j] == 1)


def test_byte_model_closed_forms():
    # 16*P*K + 8*K + 16*B per channel (K = B + 1)
    assert Bt.alg_bytes_per_block(65536, 256) == 1058824
    assert Bt.alg_bytes_per_block(131072, 256) == 2111496
    assert Bt.alg_bytes_per_block(16384, 1024) == 286984
    assert Bt.alg_bytes_per_block(40000, 256, 3) == 3 * Bt.alg_bytes_per_block(65536, 256)
    assert Bt.rounded_length(100, 256) == 256 and Bt.rounded_length(4097, 128) == 8192


def test_estimate_covers_resident_state():
    lib = T._lib.load()
    for L, B, C in [(65536, 256, 64), (4096, 128, 1), (524288, 256, 16), (131072, 256, 1024)]:
        n = S.estimate_bytes(lib, L, B, C)
        n16 = S.estimate_bytes(lib, L, B, C, T.TAN_IR_PRECISION_FP16)
        ln = Bt.rounded_length(L, B)
        # three IR slots of packed spectra (B complex per partition) and the
        # FDL, plus input/output blocks
        assert n >= C * ln * (3 * 8 + 8)
        assert n - n16 >= 3 * C * ln * 4 - 4096  # half-size IR slots
    with pytest.raises(ValueError):
        S.estimate_bytes(lib, 4096, 300, 1)
    # the partial-sum ring: 2K slots x 2 filters x chunks rows of B complex
    # per channel, so pass length and chunk cap both enter the estimate
    L, B, C = 65536, 256, 512
    base = S.estimate_bytes(lib, L, B, C, pass_blocks=4, max_tail_chunks=2)
    assert S.estimate_bytes(lib, L, B, C, pass_blocks=8, max_tail_chunks=2) - base == 4 * 4 * C * 2 * B * 8
    assert S.estimate_bytes(lib, L, B, C, pass_blocks=4, max_tail_chunks=8) - base == 4 * 4 * C * 6 * B * 8
    # two IR slots: one slot of spectra and row scales less
    assert base - S.estimate_bytes(lib, L, B, C, pass_blocks=4, max_tail_chunks=2, ir_slots=2) == \
        C * (L // B) * (B * 8 + 4)
    with pytest.raises(ValueError):
        S.estimate_bytes(lib, L, B, C, ir_slots=5)
    # uncapped, the library's own bound applies (4096 / C = 8 chunks here)
    assert S.estimate_bytes(lib, L, B, C, pass_blocks=4) == S.estimate_bytes(lib, L, B, C, pass_blocks=4,
                                                                              max_tail_chunks=8)


def test_draw_streams():
    s = Bt.draw_streams(512, seed=7)
    assert len(s) == 512
    ls = np.array([a for a, _ in s])
    assert ls.min() >= 4096 and ls.max() <= 262144
    assert set(b for _, b in s) == {128, 256, 512}
    assert s == Bt.draw_streams(512, seed=7)


def test_row_allocator_first_fit_and_coalescing():
    a = S.RowAllocator(64)
    r = [a.alloc(n) for n in (16, 8, 16, 8)]
    assert r == [0, 16, 24, 40] and a.free == [(48, 16)]
    a.release(16, 8)
    a.release(40, 8)
    assert a.free == [(16, 8), (40, 24)]
    assert a.alloc(12) == 40  # first range too short
    a.release(24, 16)  # joins both neighbours' free space: 16..40 free and 52..64
    assert a.free == [(16, 24), (52, 12)]
    with pytest.raises(RuntimeError):
        a.release(20, 2)
    assert a.alloc(100) is None
    rng = np.random.default_rng(1)
    b = S.RowAllocator(256)
    held = []
    for _ in range(3000):
        if held and rng.random() < 0.5:
            st, n = held.pop(int(rng.integers(len(held))))
            b.release(st, n)
        else:
            n = int(rng.integers(1, 17))
            st = b.alloc(n)
            if st is not None:
                assert all(st + n <= s0 or s0 + n0 <= st for s0, n0 in held)
                held.append((st, n))
        assert b.free_rows() + sum(n for _, n in held) == 256
    for st, n in held:
        b.release(st, n)
    assert b.free == [(0, 256)]


def test_nonuniform_estimate_and_options():
    lib = T._lib.load()
    NU = T.TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM
    UNI = T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
    for L, B, C in [(524288, 256, 16), (65536, 128, 4), (8192, 64, 1)]:
        u = S.estimate_bytes(lib, L, B, C, method=UNI)
        n = S.estimate_bytes(lib, L, B, C, method=NU)
        ln = Bt.rounded_length(L, B)
        # both hold three IR slots and a history of the whole IR; the frame
        # ring, input ring, partials and outputs come on top
        assert n >= C * ln * (3 * 8 + 8)
        assert n < 2 * u
        # an explicit multiple of 1 is the uniform layout
        assert S.estimate_bytes(lib, L, B, C, method=NU, nu_multiple=1) == u
    # invalid multiples: not a power of two, frame above 4096, frame above half the IR
    for mult, L, B in [(3, 65536, 256), (32, 524288, 256), (16, 4096, 256)]:
        with pytest.raises(ValueError):
            S.estimate_bytes(lib, L, B, 1, method=NU, nu_multiple=mult)


def test_nonuniform_byte_model():
    # frames cut the per-block bytes of long IRs; one frame of taps is uniform
    u = Bt.alg_bytes_per_block(524288, 256)
    n = Bt.alg_bytes_per_block(524288, 256, 1, 16)
    assert u == 16 * 2048 * 257 + 8 * 257 + 16 * 256
    k, N, pb = 257, 4096, 127
    assert n == 16 * 16 * k + 8 * k + 16 * 256 + 16 * (pb - 2) * 256 + (8 * N + 16 * N + 16 * N + 4 * N) // 16
    assert u / n > 12
    assert Bt.alg_bytes_per_block(65536, 256, 3, 1) == Bt.alg_bytes_per_block(65536, 256, 3)


def test_pass_residue_balances_pass_bytes():
    """Sessions placed one after another by pass_residue keep the pass bytes
    of the busiest tick residue within one session's pass of the quietest,
    for pass lengths 4 and 8 mixed; a session of length k lands on every
    k-th residue."""
    rng = np.random.default_rng(3)
    load = [0] * S.PASS_CYCLE
    sizes = []
    for _ in range(200):
        k = int(rng.choice([4, 8]))
        nbytes = int(rng.integers(1, 100)) * 1000
        r = S.pass_residue(load, k)
        assert 0 <= r < k
        slots = [r + j * k for j in range(S.PASS_CYCLE // k)]
        for i in slots:
            load[i] += nbytes
        sizes.append(nbytes)
    assert max(load) - min(load) <= max(sizes)
    # ties go to the lowest residue; a loaded residue is avoided
    assert S.pass_residue([0] * 8, 4) == 0
    assert S.pass_residue([5, 0, 0, 0, 0, 0, 0, 0], 8) == 1
    assert S.pass_residue([5, 0, 0, 0, 0, 0, 0, 0], 4) == 1
    assert S.pass_residue([0, 9, 9, 9, 1, 9, 9, 9], 4) == 0
    # the pass of an FP32 session: FDL and IR rows of partitions 2..P-1
    assert S.pass_bytes(65536, 256, 2) == 2 * 254 * 256 * 16
    assert S.pass_bytes(65536, 256, 2, T.TAN_IR_PRECISION_FP16) == 2 * 254 * 256 * 12


def test_layout_candidates():
    """Uniform layouts with 4- and 8-block passes for IRs of 16 or more
    partitions, one-block tails below; non-uniform multiples M = 2..16 with
    frames of at most 4,096 samples and at most half the IR."""
    assert Bt.layout_candidates(128, 4096) == [("uniform", 4), ("uniform", 8), ("nonuniform", 2),
                                               ("nonuniform", 4), ("nonuniform", 8), ("nonuniform", 16)]
    assert Bt.layout_candidates(512, 262144) == [("uniform", 4), ("uniform", 8), ("nonuniform", 2),
                                                 ("nonuniform", 4), ("nonuniform", 8)]
    assert Bt.layout_candidates(512, 4096) == [("uniform", 1), ("nonuniform", 2), ("nonuniform", 4)]
    assert Bt.layout_candidates(256, 700) == [("uniform", 1), ("nonuniform", 2)]
    for b in (128, 256, 512):
        for ln in (4096, 9000, 65536, 262144):
            for kind, v in Bt.layout_candidates(b, ln):
                if kind == "nonuniform":
                    assert v * b <= 4096 and 2 * v * b <= Bt.rounded_length(ln, b)
