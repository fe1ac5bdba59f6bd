"""Parity of the measured workloads at their stated sizes.

- The headline object: 64 channels, B = 256, 65,536 taps, a fresh IR every
  50 blocks with a one-block crossfade, processed exactly as the benchmark
  drives it (a context on the caller's stream, device memory, calls of 50
  blocks so that heads run batched beside the multi-block passes), against
  the oracle on every channel and against the FP64 crossfade model.
- An object whose streamed rows (FDL and two IR slots, 402 MB) exceed what
  the Infinity Cache can hold, so its multi-block passes read rows with
  nontemporal loads: channels 0-3 against the oracle (channels are
  independent), over several update periods and calls of uneven size.

Bar: max |gpu - oracle| <= 1e-5 of full scale and SNR >= 120 dB.
"""
import numpy as np
import pytest

import oracle_lib as O
import tan_amd as T
from helpers import crossfade_model, snr_db

pytestmark = pytest.mark.gpu
UNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
MAX_ABS = 1e-5
MIN_SNR = 120.0


def _oracle(L, B, x, updates, nb):
    """OracleConvolution over nb blocks of x (C, >= nb*B); updates maps a
    block index to the IR bank installed before it."""
    C = x.shape[0]
    orc = O.OracleConvolution(L, B, C)
    yo = np.zeros((C, nb * B), np.float32)
    xfades = []
    for b in range(nb):
        if b in updates:
            assert orc.update_td(np.ascontiguousarray(updates[b], dtype=np.float32)) == O.AMF_OK
        rc, out, _ = orc.process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]))
        assert rc == O.AMF_OK
        yo[:, b * B:(b + 1) * B] = out
        if orc.last_was_crossfade():
            xfades.append(b)
    return yo, xfades


def _check(y, yo, label):
    err = float(np.abs(y.astype(np.float64) - yo).max())
    snr = min(snr_db(yo[c], y[c]) for c in range(y.shape[0]) if np.any(yo[c]))
    print(f"{label}: max|gpu-oracle|={err:.3e} snr={snr:.1f} dB", flush=True)
    assert err <= MAX_ABS and snr >= MIN_SNR, (label, err, snr)


def test_headline_object_64ch_against_oracle():
    """The benchmark's 64-channel object over 320 blocks with 7 IR updates."""
    import torch
    L, B, C, nb, every = 65536, 256, 64, 320, 50
    hs = torch.cuda.Stream()
    ctx = T.TANContext(device=0, conv_stream=hs, general_stream=hs)
    conv = T.TANConvolution(ctx)
    assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
    rc, chunks, _ = conv.Autotune()
    assert rc == T.AMF_OK
    assert conv.Info()["pass_blocks"] == 4
    x = T.signals.input_bank(C, nb * B, base_seed=501)
    ups = {b: T.signals.ir_bank(C, L, b // every, base_seed=502) for b in range(0, nb, every)}
    print(f"headline object: {C} ch, {nb} blocks, {len(ups)} updates, {chunks} tail chunks", flush=True)
    with torch.cuda.stream(hs):
        xd = torch.from_numpy(x).cuda()
        yd = torch.zeros_like(xd)
        banks = {b: torch.from_numpy(h).cuda() for b, h in ups.items()}
        for t in range(0, nb, every):
            n = min(every, nb - t) * B
            assert conv.UpdateResponseTD(banks[t], L, None, 0) == T.AMF_OK
            rc, done = conv.Process(xd[:, t * B:], yd[:, t * B:], n)
            assert rc == T.AMF_OK and done == n
    conv.Synchronize()
    y = yd.cpu().numpy()
    crossfades = conv.Info()["crossfades"]
    conv.Terminate()
    ctx.Terminate()
    yo, xf = _oracle(L, B, x, ups, nb)
    assert xf == [b + 2 for b in sorted(ups)] and crossfades == len(xf)
    _check(y, yo, "headline 64ch")
    irs = [ups[b] for b in sorted(ups)]
    for c in (0, 37, C - 1):
        ref = crossfade_model(x[c], [i[c] for i in irs], xf, B)
        err = float(np.abs(y[c] - ref).max())
        snr = snr_db(ref, y[c])
        print(f"headline ch{c} vs FP64: max abs {err:.3e}, SNR {snr:.1f} dB", flush=True)
        assert err <= MAX_ABS and snr >= MIN_SNR


def test_streaming_pass_object_above_cache(gpu_ctx):
    """256 channels x 65,536 taps: the pass rows (402 MB) exceed the cached
    bound, so k_window's pass reads them with nontemporal loads.  Calls of
    uneven size (batched heads of 1-4 blocks beside the passes) and IR
    updates at uneven offsets, one of them issued while a switch is armed."""
    import torch
    L, B, C, nb = 65536, 256, 256, 210
    conv = T.TANConvolution(gpu_ctx)
    assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
    info = conv.Info()
    assert info["pass_blocks"] == 4 and info["cached_rows"] == 0, info
    gen = torch.Generator(device="cuda")
    gen.manual_seed(611)
    env = torch.exp(-torch.arange(L, device="cuda", dtype=torch.float32) / 12000.0)
    xd = torch.empty(C, nb * B, device="cuda").uniform_(-1.0, 1.0, generator=gen)
    yd = torch.zeros_like(xd)
    upd_blocks = [0, 37, 90, 91, 150]  # 91: the switch of 90 is armed then
    banks = {}
    for b in upd_blocks:
        h = torch.randn(C, L, device="cuda", generator=gen) * env
        banks[b] = h / h.norm(dim=1, keepdim=True)
    sizes = [1, 3, 7, 2, 13, 4, 1, 1, 9, 5]
    cuts = sorted(set(upd_blocks + [nb]))
    t, k = 0, 0
    print(f"streaming pass: {C} ch, {nb} blocks, updates {upd_blocks}", flush=True)
    while t < nb:
        if t in banks:
            assert conv.UpdateResponseTD(banks[t], L, None, 0) == T.AMF_OK
        nxt = min(c for c in cuts if c > t)
        n = min(sizes[k % len(sizes)], nxt - t)
        k += 1
        rc, done = conv.Process(xd[:, t * B:], yd[:, t * B:], n * B)
        assert rc == T.AMF_OK and done == n * B
        t += n
    conv.Synchronize()
    assert conv.Info()["cached_rows"] == 0
    conv.Terminate()
    sel = slice(0, 4)
    x = xd[sel].cpu().numpy()
    y = yd[sel].cpu().numpy()
    ups = {b: banks[b][sel].cpu().numpy() for b in upd_blocks}
    del xd, yd, banks
    torch.cuda.empty_cache()
    yo, xf = _oracle(L, B, x, ups, nb)
    print(f"streaming pass: crossfades at {xf}", flush=True)
    _check(y, yo, "streaming pass ch0-3")
