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


def test_queue_launches_equal_plain_launches():
    """Work-queue group launches (tanx_ctx_set_group_queue: the device's
    resident workgroups take the launch's items from a counter) give the
    plain group launches' outputs bit for bit: objects with multi-block
    passes of 4, 8 and 12 blocks on several phases, a non-uniform object and
    IR switches with their crossfades, over uneven calls; long enough that
    every kind (heads only, lean and crossfade passes, general) runs."""
    import torch
    NU = T.TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM
    specs = [(65536, 256, 8, UNIFORM, 12), (8000, 128, 3, UNIFORM, 0), (65536, 512, 2, UNIFORM, 8),
             (32768, 256, 4, NU, 0), (16384, 64, 5, UNIFORM, 4), (65536, 256, 16, UNIFORM, 12)]
    nb = 96
    outs = []
    for q in (False, True):
        ctx = T.TANContext(device=0)
        assert ctx.SetGroupQueue(q) == T.AMF_OK
        objs, xs, ys, irs = [], [], [], []
        for i, (L, B, C, m, kp) in enumerate(specs):
            kw = {"pass_blocks": kp, "pass_phase": (3 * i) % kp} if kp else {}
            c = T.TANConvolution(ctx, **kw)
            assert c.InitGpu(m, L, B, C) == T.AMF_OK
            irs.append([T.signals.ir_bank(C, L, u, base_seed=300 + i) for u in range(2)])
            assert c.UpdateResponseTD(irs[i][0], L, None, BLOCKING) == T.AMF_OK
            objs.append(c)
            xs.append(torch.from_numpy(T.signals.input_bank(C, nb * B, base_seed=400 + i)).cuda())
            ys.append(torch.zeros(C, nb * B, device="cuda"))
        torch.cuda.synchronize()
        blk = [0] * len(specs)
        for t in range(nb):
            idx = [i for i in range(len(specs)) if t % 3 or i % 2 == 0]
            if t % 17 == 5:
                for i in (0, 4, 5):
                    L = specs[i][0]
                    assert objs[i].UpdateResponseTD(irs[i][(t // 17) % 2 ^ 1], L, None, BLOCKING) == T.AMF_OK
            rc = T.GroupProcess([objs[i] for i in idx],
                                [xs[i][:, blk[i] * specs[i][1]:] for i in idx],
                                [ys[i][:, blk[i] * specs[i][1]:] for i in idx])
            assert rc == T.AMF_OK
            for i in idx:
                blk[i] += 1
        for c in objs:
            c.Synchronize()
        outs.append([y[:, :blk[i] * specs[i][1]].cpu().numpy() for i, y in enumerate(ys)])
        for c in objs:
            c.Terminate()
        ctx.Terminate()
    for i in range(len(specs)):
        np.testing.assert_array_equal(outs[1][i], outs[0][i], err_msg=f"object {i} {specs[i]}")
        assert float(np.abs(outs[1][i]).max()) > 0.01, i
