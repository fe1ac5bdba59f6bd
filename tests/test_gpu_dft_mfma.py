"""Block DFT on the matrix cores (precision track): a real DFT as a bf16 GEMM
with FP32 accumulation, plain and with split-bf16 compensation.  Pinned to
the FP64 DFT (numpy) and to the FP32 FFT path on the same inputs; the
compensated products must gain more than 40 dB over the plain ones."""
import numpy as np
import pytest

import tan_amd as T
from helpers import snr_db

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("log2n,C", [(9, 1024), (5, 3), (8, 17), (11, 40)])
def test_mfma_dft_against_fp64(gpu_ctx, log2n, C):
    import torch
    n = 1 << log2n
    rng = np.random.default_rng(log2n * 100 + C)
    x = rng.uniform(-1, 1, (C, n)).astype(np.float32)
    ref = np.fft.rfft(x.astype(np.float64), axis=1)  # (C, n/2 + 1)
    refri = np.stack([ref.real, ref.imag], axis=2).reshape(C, n + 2)
    fft = T.TANFFT(gpu_ctx)
    assert fft.Init() == T.AMF_OK
    xd = torch.from_numpy(x).cuda()
    got = {}
    for split in (1, 3):
        yd = torch.full((C, 2 * n), 7.0, device="cuda")
        assert fft.TransformR2CMatrixCores(log2n, C, xd, yd, split) == T.AMF_OK
        torch.cuda.synchronize()
        y = yd.cpu().numpy()
        assert np.all(y[:, n + 2:] == 7.0)  # nothing written past bin n/2
        got[split] = y[:, :n + 2]
    s1 = snr_db(refri, got[1])
    s3 = snr_db(refri, got[3])
    # bf16 operands: ~50 dB; compensated: FP32-level products
    assert s1 > 40.0 and s3 > 95.0 and s3 > s1 + 40.0, (s1, s3)
    # the FP32 FFT path on the same input (TANFFT R2C)
    xin = np.zeros((C, 2 * n), np.float32)  # TANFFT buffers hold 2n floats per channel
    xin[:, :n] = x
    yf = np.zeros((C, 2 * n), np.float32)
    rc = fft.Transform(T.TAN_FFT_R2C_TRANSFORM_DIRECTION_FORWARD, log2n, C, xin, yf)
    assert rc == T.AMF_OK
    assert snr_db(yf[:, :n + 2], got[3]) > 95.0
    assert fft.TransformR2CMatrixCores(4, C, xd, yd, 3) == T.AMF_INVALID_ARG
    assert fft.TransformR2CMatrixCores(log2n, C, xd, yd, 2) == T.AMF_INVALID_ARG
