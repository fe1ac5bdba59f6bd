"""TANFFT on the GPU against the FP64 DFT and the oracle's TANFFT layouts
(FFTImpl.cpp:507-575, 876-1122): directions 0-5, 1/N backward scaling,
planar im plane at N/2+8, bytes outside the defined ranges untouched."""
import numpy as np
import pytest

import oracle_lib as O
import tan_amd as T

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(7)


@pytest.fixture(scope="module")
def fft(gpu_ctx):
    f = T.TANFFT(gpu_ctx)
    assert f.Init() == T.AMF_OK
    return f


@pytest.mark.parametrize("log2n", range(1, 12))
def test_complex_forward_backward(fft, log2n):
    n, C = 1 << log2n, 3
    z = (RNG.standard_normal((C, n)) + 1j * RNG.standard_normal((C, n))).astype(np.complex64)
    buf = np.zeros((C, 2 * n), np.float32)
    buf[:, 0::2], buf[:, 1::2] = z.real, z.imag
    out = np.zeros_like(buf)
    assert fft.Transform(0, log2n, C, buf, out) == T.AMF_OK
    for c in range(C):
        ref = O.dft_f64(z[c].astype(np.complex128))
        got = out[c, 0::2] + 1j * out[c, 1::2]
        assert np.abs(got - ref).max() <= 3e-7 * np.abs(ref).max() * log2n
    back = np.zeros_like(buf)
    assert fft.Transform(1, log2n, C, out, back) == T.AMF_OK
    np.testing.assert_allclose(back, buf, atol=1.5e-6 * log2n)
    rc, oo = O.fft_transform(0, log2n, buf)
    assert np.abs(oo - out).max() <= 1e-6 * np.abs(oo).max() * log2n


@pytest.mark.parametrize("log2n", range(4, 13))
def test_real_transforms(fft, log2n):
    n, m, C = 1 << log2n, 1 << (log2n - 1), 2
    x = RNG.standard_normal((C, n)).astype(np.float32)
    buf = np.zeros((C, 2 * n), np.float32)
    buf[:, :n] = x
    out = np.full_like(buf, 9.0)
    assert fft.Transform(2, log2n, C, buf, out) == T.AMF_OK
    outp = np.full_like(buf, 9.0)
    assert fft.Transform(4, log2n, C, buf, outp) == T.AMF_OK
    for c in range(C):
        ref = O.dft_f64(x[c].astype(np.complex128))[:m + 1]
        got = out[c, 0:n + 2:2] + 1j * out[c, 1:n + 2:2]
        assert np.abs(got - ref).max() <= 3e-7 * np.abs(ref).max() * log2n
        gp = outp[c, :m + 1] + 1j * outp[c, m + 8:2 * m + 9]
        assert np.abs(gp - ref).max() <= 3e-7 * np.abs(ref).max() * log2n
    assert np.all(out[:, n + 2:] == 9.0)
    assert np.all(outp[:, m + 1:m + 8] == 9.0) and np.all(outp[:, 2 * m + 9:] == 9.0)
    back = np.zeros_like(buf)
    assert fft.Transform(3, log2n, C, out, back) == T.AMF_OK
    np.testing.assert_allclose(back[:, :n], x, atol=1e-6 * log2n)
    backp = np.zeros_like(buf)
    assert fft.Transform(5, log2n, C, outp, backp) == T.AMF_OK
    np.testing.assert_allclose(backp[:, :n], x, atol=1e-6 * log2n)


def test_fft_device_and_errors(fft, gpu_ctx):
    import torch
    log2n, n, C = 9, 512, 4
    x = RNG.standard_normal((C, 2 * n)).astype(np.float32)
    host = np.zeros_like(x)
    assert fft.Transform(0, log2n, C, x, host) == T.AMF_OK
    dx = torch.from_numpy(x).cuda()
    torch.cuda.synchronize()
    assert fft.Transform(0, log2n, C, dx, dx) == T.AMF_OK  # in place
    np.testing.assert_array_equal(dx.cpu().numpy(), host)
    buf = np.zeros((1, 32), np.float32)
    assert fft.Transform(4, 3, 1, buf, buf) == T.AMF_INVALID_ARG
    assert fft.Transform(6, 4, 1, buf, buf) == T.AMF_INVALID_ARG
    assert fft.Transform(0, 0, 1, buf, buf) == T.AMF_INVALID_ARG
