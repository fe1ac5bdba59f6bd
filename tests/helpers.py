"""Shared numerics for the parity tests."""
import numpy as np


def snr_db(ref, got):
    ref = np.asarray(ref, np.float64)
    err = np.asarray(got, np.float64) - ref
    e = np.sum(err * err)
    s = np.sum(ref * ref)
    if e == 0.0:
        return np.inf
    return 10.0 * np.log10(s / e)


def fft_conv_f64(x, h, n=None):
    """Linear convolution in float64 via FFT, first n samples."""
    x = np.asarray(x, np.float64)
    h = np.asarray(h, np.float64)
    n = x.size if n is None else n
    m = x.size + h.size - 1
    nfft = 1 << (m - 1).bit_length()
    y = np.fft.irfft(np.fft.rfft(x, nfft) * np.fft.rfft(h, nfft), nfft)
    return y[:n]


def crossfade_model(x, irs, switch_blocks, B, conv=fft_conv_f64):
    """FP64 model of the intended time-varying output.

    irs[k] becomes current at block switch_blocks[k] with a one-block linear
    fade from irs[k-1] (irs[-1] before the first switch is the zero IR):
    y = y_new * j/B + y_old * (B-j)/B inside the fade block."""
    n = x.size
    ys = [conv(x, h, n) for h in irs]
    y = np.zeros(n)
    nb = n // B
    ramp = np.arange(B) / B
    for b in range(nb):
        sl = slice(b * B, (b + 1) * B)
        k = -1
        for i, s in enumerate(switch_blocks):
            if b >= s:
                k = i
        if k < 0:
            continue
        if b == switch_blocks[k]:
            old = ys[k - 1][sl] if k > 0 else 0.0
            y[sl] = ys[k][sl] * ramp + old * (1.0 - ramp)
        else:
            y[sl] = ys[k][sl]
    return y
