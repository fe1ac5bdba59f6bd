"""Batched streams of uneven length and block size: one group launch per base
tick gives every stream the output of its object processed alone, and the
bucket tuner's chunk choices keep results within the FP32 tolerance (a
different chunk count only reorders the partial sums)."""
import numpy as np
import pytest

import tan_amd as T
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


def _alone(ctx, spec, x, ir):
    """The stream's object alone, one block per call as in the batch (a call
    of several blocks would run its heads together, with other rounding)."""
    import torch
    L, B, C = spec
    conv = T.TANConvolution(ctx)
    assert conv.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
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
        assert 1 <= c <= max(1, ln // b - 2)
    # a fresh batch of the same streams with the tuned chunks, from block 0
    tuned = Bt.StreamBatch(gpu_ctx, specs)
    for k, c in chosen.items():
        tuned.set_bucket_chunks(k, c)
    xs, ys = _drive(tuned, 32, seed=2)
    for j, spec in enumerate(tuned.specs):
        ya = _alone(gpu_ctx, spec, xs[j], irs[j])
        assert np.abs(ys[j] - ya).max() <= 1e-5, f"stream {j}"
