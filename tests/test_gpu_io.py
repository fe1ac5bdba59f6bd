"""Output chain on the device: TANMixer and TANConverter (all overloads) are
bit-exact against the oracle, and convolution -> mix -> convert runs device
to device with the same result as the host chain."""
import numpy as np
import pytest

import oracle_lib as O
import tan_amd as T

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("C,n", [(1, 256), (2, 1000), (16, 512), (5, 4097)])
def test_mixer_overloads_match_oracle(gpu_ctx, C, n):
    import torch
    rng = np.random.default_rng(C * 100 + n)
    x = rng.standard_normal((C, n)).astype(np.float32)
    ref = O.mix(x)
    m = T.TANMixer(gpu_ctx)
    assert m.Init(n, C) == T.AMF_OK
    out = np.zeros(n, np.float32)
    assert m.Mix(x, out) == T.AMF_OK
    np.testing.assert_array_equal(out, ref)
    xd = torch.zeros(C, n + 64, device="cuda")
    xd[:, :n] = torch.from_numpy(x)
    yd = torch.zeros(n, device="cuda")
    assert m.Mix(xd, yd) == T.AMF_OK  # strided (stride n + 64)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(yd.cpu().numpy(), ref)
    rows = [torch.from_numpy(x[c]).cuda() for c in range(C)]
    yd2 = torch.zeros(n, device="cuda")
    assert m.Mix(rows, yd2) == T.AMF_OK  # pointer list
    torch.cuda.synchronize()
    np.testing.assert_array_equal(yd2.cpu().numpy(), ref)
    assert m.Terminate() == T.AMF_OK
    assert T.TANMixer(gpu_ctx).Init(0, 2) == T.AMF_INVALID_ARG


@pytest.mark.parametrize("gain", [1.0, 0.37, 2.5])
def test_converter_matches_oracle(gpu_ctx, gain):
    import torch
    rng = np.random.default_rng(int(gain * 100))
    n = 3000
    xf = (rng.standard_normal(2 * n) * 0.6).astype(np.float32)
    xs = rng.integers(-32768, 32768, 2 * n).astype(np.int16)
    conv = T.TANConverter(gpu_ctx)
    assert conv.Init() == T.AMF_OK
    for inp in (xf, xs):
        for out_dtype in (np.float32, np.int16):
            for (ist, ioff, ost, ooff) in [(1, 0, 1, 0), (2, 1, 2, 0), (1, 5, 2, 1)]:
                m = (inp.size - ioff - 1) // ist + 1
                m = min(m, n)
                ref = np.full(2 * n + 2, 3, out_dtype)
                clip_ref = O.convert(inp, ist, ioff, ref, ost, ooff, m, gain)
                # host overload
                host = np.full(2 * n + 2, 3, out_dtype)
                rc, clipped = conv.Convert(inp, ist, m, host, ost, gain, ioff, ooff)
                assert rc == T.AMF_OK
                np.testing.assert_array_equal(host, ref)
                assert clipped == (clip_ref and inp.dtype == np.float32 and out_dtype == np.int16)
                # device overload
                di = torch.from_numpy(inp).cuda()
                do = torch.from_numpy(np.full(2 * n + 2, 3, out_dtype)).cuda()
                rc, clipped = conv.Convert(di, ist, m, do, ost, gain, ioff, ooff)
                assert rc == T.AMF_OK
                torch.cuda.synchronize()
                np.testing.assert_array_equal(do.cpu().numpy(), ref)
                assert clipped == (clip_ref and inp.dtype == np.float32 and out_dtype == np.int16)
    # errors as the reference checks them (ConverterImpl.cpp:225-228)
    assert conv.Convert(xf, 1, 0, np.zeros(4, np.int16), 1, 1.0)[0] == T.AMF_INVALID_ARG
    assert conv.Convert(xf, 1, 4, np.zeros(4, np.int16), 1, 0.0)[0] == T.AMF_INVALID_ARG


def test_convolution_mix_convert_chain_on_device(gpu_ctx):
    """SimpleVRaudio.cpp:1080-1118: device convolution output, even channels
    mixed to the left ear and odd to the right, converted to interleaved
    shorts; equals the oracle chain on the same convolution output."""
    import torch
    C, B, L, nb = 8, 256, 8192, 12
    x = T.signals.input_bank(C, nb * B, base_seed=5)
    ir = (T.signals.ir_bank(C, L, 0, base_seed=6) * 0.5).astype(np.float32)
    conv = T.TANConvolution(gpu_ctx)
    assert conv.InitGpu(T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED, L, B, C) == T.AMF_OK
    assert conv.UpdateResponseTD(ir, L, None, T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY) == T.AMF_OK
    mixer = T.TANMixer(gpu_ctx)
    assert mixer.Init(B, C // 2) == T.AMF_OK
    cvt = T.TANConverter(gpu_ctx)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros(C, B, device="cuda")
    ears = [torch.zeros(B, device="cuda") for _ in range(2)]
    pcm = torch.zeros(nb * 2 * B, dtype=torch.int16, device="cuda")
    ys = []
    for b in range(nb):
        rc, done = conv.Process(xd[:, b * B:], yd, B)
        assert rc == T.AMF_OK and done == B
        for e in range(2):
            assert mixer.Mix([yd[c] for c in range(e, C, 2)], ears[e]) == T.AMF_OK
            rc, _ = cvt.Convert(ears[e], 1, B, pcm, 2, 1.0, 0, b * 2 * B + e, want_clipped=False)
            assert rc == T.AMF_OK
        conv.Synchronize()
        ys.append(yd.cpu().numpy().copy())
    torch.cuda.synchronize()
    got = pcm.cpu().numpy()
    ref = np.zeros(nb * 2 * B, np.int16)
    for b, y in enumerate(ys):
        for e in range(2):
            O.convert(O.mix(y[e::2]), 1, 0, ref, 2, b * 2 * B + e, B, 1.0)
    np.testing.assert_array_equal(got, ref)
    assert np.abs(got).max() > 1000
