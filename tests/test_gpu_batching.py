"""Batched streams of uneven length and block size: one group launch per base
tick gives every stream the output of its object processed alone, and the
bucket tuner's chunk choices keep results within the FP32 tolerance (a
different chunk count only reorders the partial sums).  At the stated batch
size (512 streams, 4,096-262,144 taps, B in {128, 256, 512}) sampled streams
of every block size and length bucket also match the oracle over 2P blocks."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import tan_amd as T
from helpers import snr_db
from tan_amd import batching as Bt

pytestmark = pytest.mark.gpu
UNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED


def _drive(batch, nticks, seed):
    """Fresh input into every due stream before each base tick; collect outputs."""
    import torch
    rng = np.random.default_rng(seed)
    xs = [[] for _ in batch.specs]
    ys = [[] for _ in batch.specs]
    for _ in range(nticks):
        due = batch.due[batch.ticks % batch.cycle]
        for j in due:
            C, B = batch.specs[j][2], batch.specs[j][1]
            x = rng.uniform(-1, 1, (C, B)).astype(np.float32)
            batch.inputs[j].copy_(torch.from_numpy(x))
            xs[j].append(x)
        batch.tick()
        torch.cuda.synchronize()
        for j in due:
            ys[j].append(batch.outputs[j].cpu().numpy().copy())
    return [np.concatenate(x, 1) for x in xs], [np.concatenate(y, 1) for y in ys]


def _alone(ctx, spec, x, ir, method=UNIFORM, kw=None):
    """The stream's object alone (created as the batch created it), one block
    per call as in the batch (a call of several blocks would run its heads
    together, with other rounding)."""
    import torch
    L, B, C = spec
    conv = T.TANConvolution(ctx, **(kw or {}))
    assert conv.InitGpu(method, L, B, C) == T.AMF_OK
    assert conv.UpdateResponseTD(ir, L, None, T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY) == T.AMF_OK
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    for t in range(x.shape[1] // B):
        rc, done = conv.Process(xd[:, t * B:], yd[:, t * B:], B)
        assert rc == T.AMF_OK and done == B
    conv.Synchronize()
    y = yd.cpu().numpy()
    conv.Terminate()
    return y


def _irs(batch):
    """The IRs the batch installed, regenerated with its device generator."""
    import torch
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)
    out = []
    for (L, B, C) in batch.specs:
        t = torch.arange(L, device="cuda", dtype=torch.float32)
        h = torch.randn(C, L, device="cuda", generator=gen) * torch.exp(-t / (0.25 * 48000))
        out.append(h / h.norm(dim=1, keepdim=True))
        torch.empty(C, B, device="cuda").uniform_(-1.0, 1.0, generator=gen)
    return out


def test_batch_equals_alone(gpu_ctx):
    specs = [(5000, 128, 1), (40000, 256, 2), (4096, 512, 1), (70000, 128, 1), (9000, 512, 3),
             (262144, 256, 1), (20000, 512, 1), (6000, 256, 1), (131072, 128, 2), (8192, 512, 1)]
    batch = Bt.StreamBatch(gpu_ctx, specs)
    assert batch.base == 128 and batch.cycle == 4
    irs = _irs(batch)
    xs, ys = _drive(batch, 64, seed=1)
    for j, spec in enumerate(batch.specs):
        assert xs[j].shape[1] == 64 * 128
        np.testing.assert_array_equal(ys[j], _alone(gpu_ctx, spec, xs[j], irs[j]), err_msg=f"stream {j}")


def test_tuned_batch_within_tolerance(gpu_ctx):
    specs = Bt.draw_streams(48, seed=3, max_len=65536)
    batch = Bt.StreamBatch(gpu_ctx, specs)
    irs = _irs(batch)
    chosen = batch.tune(cycles=2, candidates=(1, 4, 16))
    assert set(chosen) == set(batch.buckets)
    for (b, ln), c in chosen.items():
        assert c == 0 or 1 <= c <= max(1, ln // b - 2)  # 0: the library's choice kept
    # a fresh batch of the same streams with the tuned chunks, from block 0
    tuned = Bt.StreamBatch(gpu_ctx, specs)
    for k, c in chosen.items():
        tuned.set_bucket_chunks(k, c)
    xs, ys = _drive(tuned, 32, seed=2)
    for j, spec in enumerate(tuned.specs):
        ya = _alone(gpu_ctx, spec, xs[j], irs[j])
        assert np.abs(ys[j] - ya).max() <= 1e-5, f"stream {j}"


def test_layout_tuner_streams_equal_alone(gpu_ctx):
    """The partition-layout tuner (uniform with 4- or 8-block passes, or
    non-uniform frames of M x B, per length bucket, by measured tick time)
    picks a candidate or keeps the library's choice for every bucket; rebuilt from block 0 with the chosen
    layouts, every stream equals its object alone created with the same
    options, bit for bit, and long ones match the oracle."""
    specs = Bt.draw_streams(24, seed=4, max_len=131072)
    batch = Bt.StreamBatch(gpu_ctx, specs)
    chosen = batch.tune_layouts(cycles=1)
    assert set(chosen) == set(batch.buckets)
    for (b, ln), lay in chosen.items():
        assert lay is None or lay in Bt.layout_candidates(b, ln), (b, ln, lay)  # None: the library's choice kept
    print("tuned layouts:", {f"{b}/{ln}": lay for (b, ln), lay in sorted(chosen.items())}, flush=True)
    batch.reset()
    xs, ys = _drive(batch, 48, seed=5)
    for j, spec in enumerate(batch.specs):
        method, kw = batch.stream_options(j)
        ya = _alone(gpu_ctx, spec, xs[j], batch.irs[j], method, kw)
        np.testing.assert_array_equal(ys[j], ya, err_msg=f"stream {j} {spec} {batch.layouts[j]}")


def _drive_resident(batch, nticks, seed):
    """nticks base ticks straight through the group entry point, every
    stream's whole input resident on the device: stream j's k-th block is
    columns [kB, (k+1)B) of xs[j], its output goes to the same columns of
    ys[j].  Returns (xs, ys) as CUDA tensors."""
    import torch
    lib = batch.convs[0]._lib
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    nblk = [len(range(batch.phases[j], nticks, batch.periods[j])) for j in range(len(batch.specs))]
    xs = [torch.empty(C, n * B, device="cuda").uniform_(-1.0, 1.0, generator=gen)
          for (L, B, C), n in zip(batch.specs, nblk)]
    ys = [torch.zeros_like(x) for x in xs]
    torch.cuda.synchronize()
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    strides = []
    for due in batch.due:
        strides.append(((sz * len(due))(*[xs[j].stride(0) for j in due]),
                        (sz * len(due))(*[ys[j].stride(0) for j in due])))
    step = [4 * batch.specs[j][1] for j in range(len(batch.specs))]
    xp = [x.data_ptr() for x in xs]
    yp = [y.data_ptr() for y in ys]
    for t in range(nticks):
        ph = t % batch.cycle
        due = batch.due[ph]
        blk = [(t - batch.phases[j]) // batch.periods[j] for j in due]
        ins = (vp * len(due))(*[xp[j] + step[j] * k for j, k in zip(due, blk)])
        outs = (vp * len(due))(*[yp[j] + step[j] * k for j, k in zip(due, blk)])
        rc = lib.tanx_group_process_device(len(due), batch.calls[ph].h, ins, strides[ph][0], outs, strides[ph][1],
                                           None)
        assert rc == T.AMF_OK, (t, rc)
    torch.cuda.synchronize()
    return xs, ys


def _sample_streams(specs):
    """One stream of every (block size, rounded IR length) bucket."""
    buckets = {}
    for j, (L, B, C) in enumerate(specs):
        buckets.setdefault((B, Bt.rounded_length(L, B)), []).append(j)
    return sorted(v[len(v) // 2] for v in buckets.values())


def test_c3_stated_size_against_alone_and_oracle(gpu_ctx):
    """C3 as stated: 512 single-channel streams, IR lengths log-uniform in
    [4,096, 262,144], blocks drawn from {128, 256, 512}, one group launch per
    128-sample tick.  Every stream equals its object run alone over the
    first 32 ticks (bit for bit), and a sample of streams covering every
    block size and length bucket (262,144 taps included) matches the oracle
    over at least 2P blocks of each."""
    import torch
    specs = Bt.draw_streams(512, seed=7)
    batch = Bt.StreamBatch(gpu_ctx, specs)
    assert batch.base == 128 and batch.cycle == 4 and len(batch.convs) == 512
    irs = _irs(batch)
    nticks = 4096 + 8
    print(f"C3: 512 streams, {nticks} ticks", flush=True)
    xs, ys = _drive_resident(batch, nticks, seed=21)
    # every stream alone, one block per call, over the first 32 ticks
    for j, spec in enumerate(batch.specs):
        nb = len(range(batch.phases[j], 32, batch.periods[j]))
        x = xs[j][:, :nb * spec[1]].cpu().numpy()
        y = ys[j][:, :nb * spec[1]].cpu().numpy()
        np.testing.assert_array_equal(y, _alone(gpu_ctx, spec, x, irs[j]), err_msg=f"stream {j} {spec}")
    # sampled streams against the oracle over their whole run (>= 2P blocks)
    sample = _sample_streams(batch.specs)
    rls = {Bt.rounded_length(batch.specs[j][0], batch.specs[j][1]) for j in sample}
    assert {b for j in sample for b in [batch.specs[j][1]]} == {128, 256, 512}
    assert 262144 in rls and len(sample) >= 16
    for j in sample:
        L, B, C = batch.specs[j]
        P = Bt.rounded_length(L, B) // B
        x = xs[j].cpu().numpy()
        y = ys[j].cpu().numpy()
        nb = x.shape[1] // B
        assert nb >= 2 * P, (j, nb, P)
        orc = O.OracleConvolution(L, B, C)
        assert orc.update_td(irs[j].cpu().numpy()) == O.AMF_OK
        yo = np.zeros_like(y)
        for b in range(nb):
            rc, out, _ = orc.process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]))
            assert rc == O.AMF_OK
            yo[:, b * B:(b + 1) * B] = out
        err = float(np.abs(y.astype(np.float64) - yo).max())
        snr = min(snr_db(yo[c], y[c]) for c in range(C))
        print(f"C3 stream {j}: L={L} B={B} P={P} blocks={nb} max|gpu-oracle|={err:.2e} snr={snr:.1f} dB",
              flush=True)
        assert err <= 1e-5 and snr >= 120.0, (j, err, snr)
    for c in batch.convs:
        c.Terminate()
    del xs, ys
    torch.cuda.empty_cache()
