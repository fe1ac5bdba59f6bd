"""The allocation-free group entry point (tanx_group_process_device_noalloc)
against the allocating one (tanx_group_process_device), both called through
the C ABI on twin sets of objects with the same IRs and inputs: every call's
outputs are bit-identical, over consecutive calls whose object count drops
(the second call smaller than the first) and rises again.  Objects with
multi-block passes, several block sizes and a non-uniform object with frame
kernels take part.  Calls beyond the scratch array's object limit are
rejected."""
import ctypes

import numpy as np
import pytest

import tan_amd as T

pytestmark = pytest.mark.gpu
UNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED
NONUNIFORM = T.TAN_CONVOLUTION_METHOD_FFT_PARTITIONED_NONUNIFORM
BLOCKING = T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY

# (IR taps, block, channels, method)
SPECS = [(65536, 256, 4, UNIFORM), (5000, 128, 2, UNIFORM), (131072, 512, 1, UNIFORM),
         (32768, 256, 2, NONUNIFORM), (4096, 64, 3, UNIFORM)]
# objects advanced by each call: all five, then a smaller call of three,
# then all again, ...
CALLS = [[0, 1, 2, 3, 4], [1, 3, 4]] * 24


def _objects(ctx, irs):
    objs = []
    for (L, B, C, m), h in zip(SPECS, irs):
        conv = T.TANConvolution(ctx)
        assert conv.InitGpu(m, L, B, C) == T.AMF_OK
        assert conv.UpdateResponseTD(h, L, None, BLOCKING) == T.AMF_OK
        objs.append(conv)
    return objs


def _call(fn, objs, idx, xs, ys, blk):
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    n = len(idx)
    hs = (vp * n)(*[objs[i].handle.value for i in idx])
    ins = (vp * n)(*[xs[i].data_ptr() + 4 * SPECS[i][1] * blk[i] for i in idx])
    outs = (vp * n)(*[ys[i].data_ptr() + 4 * SPECS[i][1] * blk[i] for i in idx])
    st = (sz * n)(*[xs[i].stride(0) for i in idx])
    return fn(n, hs, ins, st, outs, st, None)


def test_noalloc_group_matches_allocating_group():
    import torch
    lib = T._lib.load()
    ctx = T.TANContext(device=0)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(77)
    irs = []
    for L, B, C, m in SPECS:
        h = torch.randn(C, L, device="cuda", generator=gen) * torch.exp(
            -torch.arange(L, device="cuda", dtype=torch.float32) / 12000.0)
        irs.append(h / h.norm(dim=1, keepdim=True))
    nblk = [sum(1 for c in CALLS if i in c) for i in range(len(SPECS))]
    xs = [torch.empty(C, n * B, device="cuda").uniform_(-1.0, 1.0, generator=gen)
          for (L, B, C, m), n in zip(SPECS, nblk)]
    ya = [torch.zeros_like(x) for x in xs]
    yb = [torch.zeros_like(x) for x in xs]
    torch.cuda.synchronize()
    objs_a = _objects(ctx, irs)
    objs_b = _objects(ctx, irs)
    blk = [0] * len(SPECS)
    for k, idx in enumerate(CALLS):
        print(f"call {k}: objects {idx}", flush=True)
        assert _call(lib.tanx_group_process_device, objs_a, idx, xs, ya, blk) == T.AMF_OK
        assert _call(lib.tanx_group_process_device_noalloc, objs_b, idx, xs, yb, blk) == T.AMF_OK
        torch.cuda.synchronize()
        for i in idx:
            B = SPECS[i][1]
            got = yb[i][:, blk[i] * B:(blk[i] + 1) * B].cpu().numpy()
            want = ya[i][:, blk[i] * B:(blk[i] + 1) * B].cpu().numpy()
            np.testing.assert_array_equal(got, want, err_msg=f"call {k} object {i}")
            blk[i] += 1
    # the outputs carry signal once the IRs have faded in
    assert all(float(y.abs().max()) > 0.01 for y in yb)
    # beyond the scratch array's object limit: rejected before any work
    many = (ctypes.c_void_p * 1025)()
    sz = (ctypes.c_size_t * 1025)()
    assert lib.tanx_group_process_device_noalloc(1025, many, many, sz, many, sz, None) == T.AMF_INVALID_ARG
    for o in objs_a + objs_b:
        o.Terminate()
    ctx.Terminate()
