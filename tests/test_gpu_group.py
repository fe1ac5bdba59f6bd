"""Group launches: many convolution objects with their own block size, IR
length and channel count advance one block each in one kernel launch.  Each
object's output equals the same object processed alone, bit for bit (the
objects share no data), and matches the oracle."""
import numpy as np
import pytest

import oracle_lib as O
import tan_amd as T
from helpers import snr_db

pytestmark = pytest.mark.gpu
UNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
BLOCKING = T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY


def _streams(rng, n):
    """Uneven streams: B in {128, 256, 512}, IR lengths log-uniform."""
    out = []
    for i in range(n):
        B = int(rng.choice([128, 256, 512]))
        L = int(np.exp(rng.uniform(np.log(1024), np.log(32768))))
        C = int(rng.integers(1, 4))
        out.append((B, L, C))
    return out


def test_group_equals_alone_and_oracle(gpu_ctx):
    import torch
    rng = np.random.default_rng(5)
    specs = _streams(rng, 12)
    base, nticks = 128, 96
    objs, alone, xs, ys_g, ys_a, irs = [], [], [], [], [], []
    for i, (B, L, C) in enumerate(specs):
        nb = nticks * base // B
        x = T.signals.input_bank(C, nb * B, base_seed=100 + i)
        ir = T.signals.ir_bank(C, L, 0, base_seed=200 + i)
        irs.append(ir)
        for lst in (objs, alone):
            c = T.TANConvolution(gpu_ctx)
            assert c.InitGpu(UNIFORM, L, B, C) == T.AMF_OK
            assert c.UpdateResponseTD(ir, L, None, BLOCKING) == T.AMF_OK
            lst.append(c)
        xs.append(torch.from_numpy(x).cuda())
        ys_g.append(torch.zeros(C, nb * B, device="cuda"))
        ys_a.append(torch.zeros(C, nb * B, device="cuda"))
    torch.cuda.synchronize()
    for t in range(nticks):
        due = [i for i, (B, L, C) in enumerate(specs) if (t * base) % B == 0]
        blk = {i: (t * base) // specs[i][0] for i in due}
        rc = T.GroupProcess([objs[i] for i in due],
                            [xs[i][:, blk[i] * specs[i][0]:] for i in due],
                            [ys_g[i][:, blk[i] * specs[i][0]:] for i in due])
        assert rc == T.AMF_OK
        for i in due:
            B = specs[i][0]
            rc, done = alone[i].Process(xs[i][:, blk[i] * B:], ys_a[i][:, blk[i] * B:], B)
            assert rc == T.AMF_OK and done == B
    for c in objs + alone:
        c.Synchronize()
    for i, (B, L, C) in enumerate(specs):
        g = ys_g[i].cpu().numpy()
        a = ys_a[i].cpu().numpy()
        np.testing.assert_array_equal(g, a)
        orc = O.OracleConvolution(L, B, C)
        orc.update_td(irs[i])
        x = xs[i].cpu().numpy()
        yo = np.zeros_like(x)
        for b in range(x.shape[1] // B):
            _, o, _ = orc.process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]))
            yo[:, b * B:(b + 1) * B] = o
        assert np.abs(g - yo).max() <= 1e-5
        assert min(snr_db(yo[c], g[c]) for c in range(C)) >= 120.0
    for c in objs + alone:
        c.Terminate()


def test_group_rejects_mixed_objects(gpu_ctx):
    import torch
    a = T.TANConvolution(gpu_ctx)
    a.InitGpu(UNIFORM, 4096, 256, 1)
    b = T.TANConvolution(gpu_ctx, ir_precision=T.TAN_IR_PRECISION_FP16)
    b.InitGpu(UNIFORM, 4096, 256, 1)
    x = torch.zeros(1, 256, device="cuda")
    y = torch.zeros(1, 256, device="cuda")
    assert T.GroupProcess([a, b], [x, x], [y, y]) == T.AMF_NOT_SUPPORTED
    assert T.GroupProcess([a, a], [x, x], [y, y]) == T.AMF_INVALID_ARG
    a.Terminate()
    b.Terminate()
