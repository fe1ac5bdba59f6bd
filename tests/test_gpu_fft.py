"""TANFFT on the GPU against the FP64 DFT and the oracle's TANFFT layouts
(FFTImpl.cpp:507-575, 876-1122): directions 0-5, 1/N backward scaling,
planar im plane at N/2+8, bytes outside the defined ranges untouched, at
every size up to the reference's 2^19 (FFTImpl.h:67)."""
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


# Sizes above one workgroup's LDS (2,048 complex points) run the four-step
# transform; numpy's FP64 FFT is the reference there (an O(N^2) DFT is too
# slow), and closed forms pin its index mapping: a tone lands in one bin, an
# impulse gives a pure phase ramp.
LARGE_COMPLEX = [12, 13, 15, 17, 19]
LARGE_REAL = [13, 14, 16, 19]


@pytest.mark.parametrize("log2n", LARGE_COMPLEX)
def test_complex_four_step(fft, log2n):
    n, C = 1 << log2n, 2
    rng = np.random.default_rng(100 + log2n)
    z = (rng.standard_normal((C, n)) + 1j * rng.standard_normal((C, n))).astype(np.complex64)
    buf = np.zeros((C, 2 * n), np.float32)
    buf[:, 0::2], buf[:, 1::2] = z.real, z.imag
    out = np.zeros_like(buf)
    print(f"four-step complex log2n={log2n}", flush=True)
    assert fft.Transform(0, log2n, C, buf, out) == T.AMF_OK
    for c in range(C):
        ref = np.fft.fft(z[c].astype(np.complex128))
        got = out[c, 0::2] + 1j * out[c, 1::2]
        assert np.abs(got - ref).max() <= 3e-7 * np.abs(ref).max() * log2n
    back = np.zeros_like(buf)
    assert fft.Transform(1, log2n, C, out, back) == T.AMF_OK
    np.testing.assert_allclose(back, buf, atol=1.5e-6 * log2n)
    # closed forms: tone k0 -> n at bin k0; impulse at n0 -> W^(k n0)
    k0, n0 = (n // 3) | 1, (n // 5) + 7
    t = np.zeros((2, 2 * n), np.float32)
    ph = 2 * np.pi * k0 * np.arange(n) / n
    t[0, 0::2], t[0, 1::2] = np.cos(ph), np.sin(ph)
    t[1, 2 * n0] = 1.0
    to = np.zeros_like(t)
    assert fft.Transform(0, log2n, 2, t, to) == T.AMF_OK
    tone = to[0, 0::2] + 1j * to[0, 1::2]
    assert int(np.argmax(np.abs(tone))) == k0 and abs(tone[k0] - n) <= 1e-3 * n
    assert np.abs(np.delete(tone, k0)).max() <= 1e-3 * n ** 0.5
    ramp = to[1, 0::2] + 1j * to[1, 1::2]
    want = np.exp(-2j * np.pi * ((np.arange(n) * n0) % n) / n)
    assert np.abs(ramp - want).max() <= 2e-6 * log2n


@pytest.mark.parametrize("log2n", LARGE_REAL)
def test_real_four_step(fft, log2n):
    n, m, C = 1 << log2n, 1 << (log2n - 1), 2
    rng = np.random.default_rng(200 + log2n)
    x = rng.standard_normal((C, n)).astype(np.float32)
    buf = np.zeros((C, 2 * n), np.float32)
    buf[:, :n] = x
    out = np.full_like(buf, 9.0)
    outp = np.full_like(buf, 9.0)
    print(f"four-step real log2n={log2n}", flush=True)
    assert fft.Transform(2, log2n, C, buf, out) == T.AMF_OK
    assert fft.Transform(4, log2n, C, buf, outp) == T.AMF_OK
    for c in range(C):
        ref = np.fft.rfft(x[c].astype(np.float64))
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


def test_four_step_device_in_place(fft):
    """Device buffers, in place, at 2^19 points: the host path's bits."""
    import torch
    log2n, n, C = 19, 1 << 19, 3
    x = np.random.default_rng(5).standard_normal((C, 2 * n)).astype(np.float32)
    for mode in (0, 2):
        host = np.zeros_like(x)
        assert fft.Transform(mode, log2n, C, x, host) == T.AMF_OK
        dx = torch.from_numpy(x).cuda()
        torch.cuda.synchronize()
        assert fft.Transform(mode, log2n, C, dx, dx) == T.AMF_OK
        got = dx.cpu().numpy()
        defined = 2 * n if mode == 0 else n + 2
        np.testing.assert_array_equal(got[:, :defined], host[:, :defined])


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
    # the reference's largest plan is 2^19 (FFTImpl.h:67)
    big = np.zeros((1, 2 << 20), np.float32)
    assert fft.Transform(0, 20, 1, big, big) == T.AMF_NOT_SUPPORTED
    assert fft.Transform(2, 20, 1, big, big) == T.AMF_NOT_SUPPORTED
