"""IR updates of several objects in one grouped call (GroupUpdateResponseTD,
tanx_group_update_td_device) against the same updates made one object at a
time on twin objects: block sizes 64-512, the non-uniform method's frame
partitions, FP16 and BF16 spectra, and a channel held by STOP_INPUT all give
the same output bits; the mix is checked against the oracle where the
spectra are FP32."""
import numpy as np
import pytest

import oracle_lib as O
import tan_amd as T

pytestmark = pytest.mark.gpu
UNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
NONUNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM

SPECS = [  # (ir_len, block, channels, method, ir precision)
    (65536, 256, 3, UNIFORM, T.TAN_IR_PRECISION_FP32),
    (5000, 128, 2, UNIFORM, T.TAN_IR_PRECISION_FP32),
    (131072, 512, 1, UNIFORM, T.TAN_IR_PRECISION_FP32),
    (32768, 256, 2, NONUNIFORM, T.TAN_IR_PRECISION_FP32),
    (20000, 256, 2, UNIFORM, T.TAN_IR_PRECISION_FP16),
    (4096, 64, 3, UNIFORM, T.TAN_IR_PRECISION_BF16),
]


def test_grouped_updates_equal_per_object_updates(gpu_ctx):
    import torch
    nb, every = 60, 7
    objs = []
    for i, (L, B, C, m, prec) in enumerate(SPECS):
        pair = []
        for _ in range(2):
            c = T.TANConvolution(gpu_ctx, ir_precision=prec)
            assert c.InitGpu(m, L, B, C) == T.AMF_OK
            pair.append(c)
        x = T.signals.input_bank(C, nb * B, base_seed=900 + i)
        irs = [T.signals.ir_bank(C, L, k, base_seed=70 + i) for k in range(3)]
        objs.append((pair, x, torch.from_numpy(x).cuda(), [torch.from_numpy(h).cuda() for h in irs],
                     [torch.zeros(C, nb * B, device="cuda") for _ in range(2)], L, B, C))
    hold = np.zeros(SPECS[0][2], np.uint32)
    hold[1] = T.TAN_CONVOLUTION_CHANNEL_FLAG_STOP_INPUT
    ups = {}
    for t in range(nb):
        if t % every == 0:
            k = (t // every) % 3
            flags = [hold if (t and j == 0) else None for j in range(len(objs))]
            for j, (pair, _x, _xd, irs, _ys, L, _B, _C) in enumerate(objs):
                assert pair[0].UpdateResponseTD(irs[k], L, flags[j]) == T.AMF_OK
            assert T.GroupUpdateResponseTD([o[0][1] for o in objs], [o[3][k] for o in objs],
                                           [o[5] for o in objs], flags) == T.AMF_OK
            ups[t] = k
        for pair, _x, xd, _irs, ys, L, B, C in objs:
            for j in range(2):
                rc, done = pair[j].Process(xd[:, t * B:], ys[j][:, t * B:], B)
                assert rc == T.AMF_OK and done == B
    torch.cuda.synchronize()
    for i, (pair, x, _xd, irs, ys, L, B, C) in enumerate(objs):
        y0, y1 = ys[0].cpu().numpy(), ys[1].cpu().numpy()
        np.testing.assert_array_equal(y1, y0, err_msg=f"object {i} {SPECS[i]}")
        assert np.abs(y0).max() > 0.01
        if SPECS[i][3] == UNIFORM and SPECS[i][4] == T.TAN_IR_PRECISION_FP32 and i != 0:
            orc = O.OracleConvolution(L, B, C)
            yo = np.zeros_like(x)
            for b in range(nb):
                if b in ups:
                    assert orc.update_td(T.signals.ir_bank(C, L, ups[b], base_seed=70 + i)) == O.AMF_OK
                yo[:, b * B:(b + 1) * B] = orc.process(np.ascontiguousarray(x[:, b * B:(b + 1) * B]))[1]
            err = np.abs(y0.astype(np.float64) - yo).max()
            print(f"object {i}: max|gpu-oracle| = {err:.2e}", flush=True)
            assert err <= 1e-5
    for pair, *_r in objs:
        for c in pair:
            c.Terminate()


def test_grouped_update_rejections(gpu_ctx):
    import torch
    a = T.TANConvolution(gpu_ctx)
    assert a.InitGpu(UNIFORM, 4096, 256, 2) == T.AMF_OK
    h = torch.zeros(2, 4096, device="cuda")
    # the same object twice, a view shorter than the length, a time-domain object
    assert T.GroupUpdateResponseTD([a, a], [h, h]) == T.AMF_INVALID_ARG
    assert T.GroupUpdateResponseTD([a], [h[:, :100]], [4096]) == T.AMF_INVALID_ARG
    td = T.TANConvolution(gpu_ctx)
    assert td.InitGpu(T.TAN_CONVOLUTION_METHOD_TIME_DOMAIN, 256, 256, 2) == T.AMF_OK
    assert T.GroupUpdateResponseTD([a, td], [h, h[:, :256]]) == T.AMF_NOT_SUPPORTED
    a.Terminate()
    td.Terminate()
