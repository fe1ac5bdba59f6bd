"""Pins the CPU oracle (oracle/tan_oracle.c) before it is trusted.

The reference ships no expected outputs for this path (its drivers write WAV
files and never compare), and its FFT provider (FFTW 3.3.5) is absent, so the
oracle is pinned by mathematics: the FP64 DFT against numpy float64 and
closed forms, the FP64 direct convolution against numpy, the restated FFT
against the FP64 DFT, and the restated UPOLS convolution against the FP64
direct convolution, plus invariants.  Planted mistakes in the oracle must
fail these checks.
"""
import numpy as np
import pytest

import oracle_lib as O
from helpers import crossfade_model, fft_conv_f64, snr_db

RNG = np.random.default_rng(20261202)


# ---------------------------------------------------------------- FP64 layer
@pytest.mark.parametrize("n", [1, 2, 3, 8, 12, 64, 125])
def test_dft_f64_matches_numpy(n):
    x = RNG.standard_normal(n) + 1j * RNG.standard_normal(n)
    np.testing.assert_allclose(O.dft_f64(x), np.fft.fft(x), rtol=0, atol=1e-12 * n)
    np.testing.assert_allclose(O.dft_f64(x, inverse=True), np.fft.ifft(x) * n, rtol=0, atol=1e-12 * n)


def test_dft_f64_closed_forms():
    n = 32
    imp = np.zeros(n, complex)
    imp[0] = 1.0
    np.testing.assert_allclose(O.dft_f64(imp), np.ones(n), atol=1e-14)
    k = 5
    cosk = np.cos(2 * np.pi * k * np.arange(n) / n)
    X = O.dft_f64(cosk)
    exp = np.zeros(n, complex)
    exp[k] = exp[n - k] = n / 2
    np.testing.assert_allclose(X, exp, atol=1e-12)


def test_direct_conv_f64_matches_numpy():
    x = RNG.uniform(-1, 1, 1000).astype(np.float32)
    h = RNG.standard_normal(77).astype(np.float32)
    np.testing.assert_allclose(O.direct_conv_f64(x, h), np.convolve(x.astype(np.float64), h)[:1000],
                               rtol=0, atol=1e-12)
    full = O.direct_conv_f64(x, h, ny=1076)
    np.testing.assert_allclose(full, np.convolve(x.astype(np.float64), h), rtol=0, atol=1e-12)


# ---------------------------------------------------------- FFT restatement
def _cbuf(z):
    b = np.zeros((1, 2 * z.size), np.float32)
    b[0, 0::2] = z.real
    b[0, 1::2] = z.imag
    return b


@pytest.mark.parametrize("log2n", [1, 2, 3, 4, 7, 8, 10, 11])
def test_fft_complex_against_fp64_dft(log2n):
    n = 1 << log2n
    z = (RNG.standard_normal(n) + 1j * RNG.standard_normal(n)).astype(np.complex64)
    ref = O.dft_f64(z.astype(np.complex128))
    rc, out = O.fft_transform(0, log2n, _cbuf(z))
    assert rc == O.AMF_OK
    got = out[0, 0::2] + 1j * out[0, 1::2]
    assert np.abs(got - ref).max() <= 4e-7 * np.abs(ref).max() * max(1, log2n)
    # backward is scaled by 1/N (FFTImpl.cpp:906-913)
    rc, back = O.fft_transform(1, log2n, out)
    zb = back[0, 0::2] + 1j * back[0, 1::2]
    assert np.abs(zb - z).max() <= 1e-6 * max(1, log2n)


@pytest.mark.parametrize("log2n", [4, 5, 8, 9, 11, 12])
def test_fft_real_layouts_against_fp64_dft(log2n):
    n, m = 1 << log2n, 1 << (log2n - 1)
    x = RNG.standard_normal(n).astype(np.float32)
    ref = O.dft_f64(x.astype(np.complex128))[: m + 1]
    buf = np.zeros((1, 2 * n), np.float32)
    buf[0, :n] = x
    rc, out = O.fft_transform(2, log2n, buf)  # interleaved
    assert rc == O.AMF_OK
    got = out[0, 0:n + 2:2] + 1j * out[0, 1:n + 2:2]
    assert np.abs(got - ref).max() <= 5e-7 * np.abs(ref).max() * log2n
    rc, outp = O.fft_transform(4, log2n, buf)  # planar: im at N/2 + 8
    gp = outp[0, :m + 1] + 1j * outp[0, m + 8:2 * m + 9]
    assert np.abs(gp - ref).max() <= 5e-7 * np.abs(ref).max() * log2n
    rc, back = O.fft_transform(5, log2n, outp)
    assert np.abs(back[0, :n] - x).max() <= 2e-6 * log2n
    rc, back2 = O.fft_transform(3, log2n, out)
    assert np.abs(back2[0, :n] - x).max() <= 2e-6 * log2n


def test_fft_closed_forms_and_parseval():
    log2n, n = 9, 512
    buf = np.zeros((1, 2 * n), np.float32)
    buf[0, 0] = 1.0
    rc, out = O.fft_transform(2, log2n, buf)
    np.testing.assert_allclose(out[0, 0:n + 2:2], 1.0, atol=1e-7)
    np.testing.assert_allclose(out[0, 1:n + 2:2], 0.0, atol=1e-7)
    k = 37
    buf[0, :n] = np.cos(2 * np.pi * k * np.arange(n) / n)
    rc, out = O.fft_transform(2, log2n, buf)
    X = out[0, 0:n + 2:2] + 1j * out[0, 1:n + 2:2]
    exp = np.zeros(n // 2 + 1, complex)
    exp[k] = n / 2
    assert np.abs(X - exp).max() < 2e-4
    x = RNG.standard_normal(n).astype(np.float32)
    buf[0, :n] = x
    rc, out = O.fft_transform(2, log2n, buf)
    X = out[0, 0:n + 2:2] + 1j * out[0, 1:n + 2:2]
    e_freq = (np.abs(X[0]) ** 2 + np.abs(X[-1]) ** 2 + 2 * np.sum(np.abs(X[1:-1]) ** 2)) / n
    assert abs(e_freq - np.sum(x.astype(np.float64) ** 2)) < 1e-5 * e_freq


def test_fft_invalid_args():
    buf = np.zeros((1, 16), np.float32)
    assert O.fft_transform(4, 3, buf)[0] == O.AMF_INVALID_ARG  # planar needs N >= 16
    assert O.fft_transform(7, 3, buf)[0] == O.AMF_INVALID_ARG
    assert O.fft_transform(0, 0, buf)[0] == O.AMF_INVALID_ARG


# ----------------------------------------------------- UPOLS restatement
def _run(conv, x, B, flags=None):
    C, n = x.shape
    y = np.zeros_like(x)
    xf = []
    for b in range(n // B):
        rc, out, done = conv.process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]), flags=flags)
        assert rc == O.AMF_OK and done == B
        y[:, b * B:(b + 1) * B] = out
        xf.append(conv.last_was_crossfade())
    return y, xf


def _decaying_ir(C, L, tau=12000.0, seed=1):
    r = np.random.default_rng(seed)
    h = r.standard_normal((C, L)) * np.exp(-np.arange(L) / tau)
    h /= np.sqrt(np.sum(h * h, axis=1, keepdims=True))
    return h.astype(np.float32)


def test_upols_c1_against_fp64_direct():
    """C1: 2 ch, B=1024, L=16384, 200 blocks, IR installed before block 0:
    blocks 0 and 1 are silent, block 2 fades in (the reference's schedule)."""
    L, B, C, nb = 16384, 1024, 2, 200
    h = _decaying_ir(C, L)
    x = RNG.uniform(-1, 1, (C, nb * B)).astype(np.float32)
    conv = O.OracleConvolution(L, B, C)
    assert conv.update_td(h) == O.AMF_OK
    y, xf = _run(conv, x, B)
    assert [i for i, v in enumerate(xf) if v] == [2]
    for c in range(C):
        ref = crossfade_model(x[c], [h[c]], [2], B, conv=lambda a, b, n: O.direct_conv_f64(a, b, n))
        assert np.all(y[c, :2 * B] == 0.0)
        assert np.abs(y[c] - ref).max() <= 1e-5
        assert snr_db(ref, y[c]) >= 120.0


def test_upols_updates_crossfade_model():
    L, B, C, nb = 2048, 128, 3, 60
    irs = [_decaying_ir(C, L, tau=400.0, seed=s) for s in range(4)]
    x = RNG.uniform(-1, 1, (C, nb * B)).astype(np.float32)
    conv = O.OracleConvolution(L, B, C)
    y = np.zeros_like(x)
    upd_at = [0, 10, 25, 40]
    xfs = []
    for b in range(nb):
        if b in upd_at:
            assert conv.update_td(irs[upd_at.index(b)]) == O.AMF_OK
        rc, out, _ = conv.process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]))
        y[:, b * B:(b + 1) * B] = out
        if conv.last_was_crossfade():
            xfs.append(b)
    assert xfs == [u + 2 for u in upd_at]
    for c in range(C):
        ref = crossfade_model(x[c], [i[c] for i in irs], xfs, B)
        assert np.abs(y[c] - ref).max() <= 1e-5
        assert snr_db(ref, y[c]) >= 120.0


def test_upols_finalize_mode_matches():
    L, B, C, nb = 2048, 256, 2, 30
    h1, h2 = _decaying_ir(C, L, 500, 3), _decaying_ir(C, L, 500, 4)
    x = RNG.uniform(-1, 1, (C, nb * B)).astype(np.float32)
    outs = []
    for fmode in (False, True):
        conv = O.OracleConvolution(L, B, C, finalize_mode=fmode)
        y = np.zeros_like(x)
        for b in range(nb):
            if b in (0, 12):
                conv.update_td(h1 if b == 0 else h2)
            rc, out, _ = conv.process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]))
            if fmode:
                conv.finalize()
            y[:, b * B:(b + 1) * B] = out
        outs.append(y)
    np.testing.assert_array_equal(outs[0], outs[1])


def test_reference_crossfade_defect_window():
    """The reference's crossfade block advances the FDL twice
    (ConvolutionImpl.cpp:976-991); its output deviates inside [s, s+P+1] and
    equals the intended output outside it."""
    L, B, C, nb = 2048, 256, 1, 40
    P = L // B
    h1, h2 = _decaying_ir(C, L, 500, 5), _decaying_ir(C, L, 500, 6)
    x = RNG.uniform(-1, 1, (C, nb * B)).astype(np.float32)
    ys = {}
    for mode in (O.XFADE_INTENDED, O.XFADE_REFERENCE):
        conv = O.OracleConvolution(L, B, C, xfade=mode)
        y = np.zeros_like(x)
        for b in range(nb):
            if b in (0, 15):
                conv.update_td(h1 if b == 0 else h2)
            rc, out, _ = conv.process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]))
            y[:, b * B:(b + 1) * B] = out
        ys[mode] = y
    d = np.abs(ys[0] - ys[1]).reshape(C, nb, B).max(axis=(0, 2))
    for s in (2, 17):
        assert d[s:s + P + 2].max() > 1e-3
    outside = [b for b in range(nb) if not any(s <= b <= s + P + 1 for s in (2, 17))]
    assert d[outside].max() <= 1e-5


def test_upols_invariants():
    L, B, C, nb = 1024, 64, 3, 24
    h = _decaying_ir(C, L, 200, 9)
    # impulse on channel 0 reproduces h (after the fade-in at block 2)
    x = np.zeros((C, nb * B), np.float32)
    x[0, 3 * B] = 1.0
    x[1] = RNG.uniform(-1, 1, nb * B)
    conv = O.OracleConvolution(L, B, C)
    conv.update_td(h)
    y, _ = _run(conv, x, B)
    np.testing.assert_allclose(y[0, 3 * B:3 * B + L], h[0], atol=2e-7)
    assert np.all(y[2] == 0.0)  # zero input, any IR
    # linearity in x
    conv2 = O.OracleConvolution(L, B, C)
    conv2.update_td(h)
    y2, _ = _run(conv2, 2.0 * x, B)
    np.testing.assert_allclose(y2, 2.0 * y, atol=1e-6)
    # zero IR -> zero output
    conv3 = O.OracleConvolution(L, B, C)
    conv3.update_td(np.zeros((C, L), np.float32))
    y3, _ = _run(conv3, x, B)
    assert np.all(y3 == 0.0)


def test_audiobuffer_kat():
    """KAT built as the reference's test helpers do: pulse-train sine bursts
    (AudioBuffer.cpp:255-279) convolved with the 3-tap IR {1, 0.75, 0.5} at
    offsets {0, I/4, I/2} (AudioBuffer.cpp:290-307); closed form
    y[n] = x[n] + 0.75 x[n - I/4] + 0.5 x[n - I/2]."""
    interval, B, C, nb = 512, 128, 2, 40
    n = np.arange(nb * B)
    pw = interval >> 2
    pd = pw >> 2
    freq = np.float32(1.0 / pd)
    x = np.where(n % interval < pw, 0.5 * np.sin(n.astype(np.float32) * 2.0 * np.float32(np.pi) * freq), 0.0)
    x = np.tile(x.astype(np.float32), (C, 1))
    h = np.zeros((C, interval), np.float32)
    h[:, 0], h[:, interval >> 2], h[:, interval >> 1] = 1.0, 0.75, 0.5
    conv = O.OracleConvolution(interval, B, C)
    conv.update_td(h)
    y, _ = _run(conv, x, B)
    xd = x[0].astype(np.float64)
    ref = xd.copy()
    ref[pw:] += 0.75 * xd[:-pw]
    ref[2 * pw:] += 0.5 * xd[:-2 * pw]
    ramp = np.concatenate([np.zeros(2 * B), np.arange(B) / B, np.ones(nb * B - 3 * B)])
    ref = ref * ramp
    assert np.abs(y[0] - ref).max() / np.finfo(np.float32).eps < 64


def test_channel_flags_and_free_channel():
    L, B, C = 512, 64, 3
    h = _decaying_ir(C, L, 100, 11)
    conv = O.OracleConvolution(L, B, C)
    conv.update_td(h)
    x = RNG.uniform(-1, 1, (C, B)).astype(np.float32)
    for _ in range(4):
        conv.process(x)
    stop = np.array([0, 1, 0], np.uint32)  # STOP_INPUT on channel 1
    rc, idx = conv.next_free_channel(stop)
    assert rc == O.AMF_NOT_FOUND
    # channel 1 drains for m_length samples, then is free
    for _ in range(L // B):
        rc, out, _ = conv.process(x, flags=stop)
        assert rc == O.AMF_OK
    rc, idx = conv.next_free_channel(stop)
    assert rc == O.AMF_OK and idx == 1
    # output of a stopped (available) channel is left untouched
    sentinel = np.full((C, B), 7.0, np.float32)
    rc, out, _ = conv.process(x, out=sentinel, flags=stop)
    assert np.all(out[1] == 7.0) and not np.all(out[0] == 7.0)
    # every channel stopped and flushed -> AMF_WRONG_STATE
    allstop = np.array([3, 3, 3], np.uint32)
    rc, _, _ = conv.process(x, flags=allstop)
    assert rc == O.AMF_WRONG_STATE


def test_partial_block_not_processed():
    conv = O.OracleConvolution(256, 64, 1)
    x = np.ones((1, 32), np.float32)
    rc, out, done = conv.process(x, n=32)
    assert rc == O.AMF_OK and done == 0


def test_update_longer_than_init_length_rejected():
    conv = O.OracleConvolution(256, 64, 1)
    assert conv.update_td(np.zeros((1, 300), np.float32)) == O.AMF_INVALID_ARG


# --------------------------------------------------------- planted mistakes
@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 5])
def test_planted_mistakes_are_caught(mutation):
    """Each plausible mistake (conjugated operand, dropped partition, ring
    walked backwards, missing 1/N, swapped crossfade weights) must fail the
    FP64 check used above."""
    L, B, C, nb = 1024, 128, 1, 20
    h = _decaying_ir(C, L, 300, 13)
    x = RNG.uniform(-1, 1, (C, nb * B)).astype(np.float32)
    O.set_mutation(mutation)
    try:
        conv = O.OracleConvolution(L, B, C)
        conv.update_td(h)
        y, _ = _run(conv, x, B)
    finally:
        O.set_mutation(0)
    ref = crossfade_model(x[0], [h[0]], [2], B)
    assert np.abs(y[0] - ref).max() > 1e-3


def test_fp64_fft_conv_matches_direct():
    x = RNG.uniform(-1, 1, 5000).astype(np.float32)
    h = RNG.standard_normal(1500).astype(np.float32)
    np.testing.assert_allclose(fft_conv_f64(x, h), O.direct_conv_f64(x, h), atol=1e-10)
