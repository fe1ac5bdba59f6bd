"""Room responses on the GPU (tanx_room_*) against the oracle's restatement
of TrueAudioVR's generator: bit for bit, since both sum the same
single-precision image contributions in 2^26 fixed point; many responses in
one launch; and a generated response fed to the convolution engine on
device memory."""
import ctypes

import numpy as np
import pytest

import oracle_lib as O
import tan_amd as T
from tan_amd import room as R
from helpers import fft_conv_f64, snr_db

pytestmark = pytest.mark.gpu

CASES = [
    # (w, h, l, damp L R T B F Bk), source, head, (yaw, pitch, roll), length, flags, bounces
    ((5.0, 3.0, 4.0, 0.6, 0.7, 0.5, 0.8, 0.65, 0.55), (1.5, 1.2, 2.0), (3.0, 1.6, 2.5), (0, 0, 0), 8192, 0, 0),
    ((4.3, 2.7, 6.1, 0.9, 0.9, 0.3, 0.4, 0.8, 0.85), (0.5, 2.0, 5.0), (2.1, 1.7, 1.0), (35, -10, 5), 30000, 0, 0),
    ((3.0, 3.0, 3.0, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5), (9.0, -1.0, 1.0), (1.0, 1.5, 2.0), (90, 0, 0), 4096, 0x40, 3),
    ((6.0, 2.5, 4.5, 1.0, 0.2, 1.0, 0.0, 0.7, 0.3), (2.0, 1.0, 1.0), (4.0, 1.2, 3.5), (-120, 20, 0), 6000, 0x80, 0),
    ((12.0, 4.0, 9.0, 0.95, 0.95, 0.9, 0.9, 0.95, 0.95), (3.0, 2.0, 7.0), (8.0, 1.7, 2.0), (10, 0, 0), 65536, 0, 0),
    ((2.0, 2.2, 2.4, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0), (1.0, 1.1, 1.2), (1.0, 1.1, 1.25), (0, 0, 0), 2048, 0, 0),
]


def _args(case):
    dims, src, head, ang, n, flags, bounces = CASES[case]
    return (R.RoomDefinition(*dims), R.MonoSource(*src), R.make_listener(head, 0.16, *ang), n, flags, bounces)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_room_response_bit_exact(gpu_ctx, case):
    room, src, lst, n, flags, bounces = _args(case)
    print(f"room case {case}: length {n}", flush=True)
    left = np.zeros(n, np.float32)
    right = np.zeros(n, np.float32)
    rc = R.generateRoomResponse(gpu_ctx, room, src, lst, 48000, n, left, right, flags, bounces)
    assert rc == T.AMF_OK
    rc, oleft, oright = O.room_response(room, src, lst, 48000, n, flags, bounces)
    assert rc == 0
    np.testing.assert_array_equal(left, oleft)
    np.testing.assert_array_equal(right, oright)
    assert np.count_nonzero(left) > 0
    # the accumulator is left clear: a second call gives the same result
    rc = R.generateRoomResponse(gpu_ctx, room, src, lst, 48000, n, left, right, flags, bounces)
    np.testing.assert_array_equal(left, oleft)


def test_many_responses_in_one_launch(gpu_ctx):
    import torch
    n = 12000
    items = [_args(c) for c in (0, 1, 3, 5)]
    # the same flags for all: bounds and suppression apply to the whole launch
    rooms = [a[0] for a in items] * 2
    srcs = [a[1] for a in items] + [R.MonoSource(1.0 + 0.3 * k, 1.0, 1.5) for k in range(4)]
    lsts = [a[2] for a in items] * 2
    out = torch.full((2 * len(rooms), n + 37), -1.0, device="cuda")
    rc = R.GenerateRoomResponses(gpu_ctx, rooms, srcs, lsts, 48000, n, out)
    assert rc == T.AMF_OK
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    for k in range(len(rooms)):
        rc, ol, orr = O.room_response(rooms[k], srcs[k], lsts[k], 48000, n)
        assert rc == 0
        np.testing.assert_array_equal(got[2 * k, :n], ol, err_msg=f"response {k}")
        np.testing.assert_array_equal(got[2 * k + 1, :n], orr, err_msg=f"response {k}")
    assert (got[:, n:] == -1.0).all()  # nothing written past the length


def test_generated_response_drives_the_convolution(gpu_ctx):
    """IR source -> UpdateResponseTD on device memory -> Process: the output
    is the convolution with the generated response."""
    import torch
    L, B = 16384, 256
    room, src, lst, _, _, _ = _args(0)
    ir = torch.empty((2, L), device="cuda")
    assert R.GenerateRoomResponses(gpu_ctx, [room], [src], [lst], 48000, L, ir) == T.AMF_OK
    conv = T.TANConvolution(gpu_ctx)
    assert conv.InitGpu(T.TAN_CONVOLUTION_METHOD_FFT_UNIFORM_PARTITIONED, L, B, 2) == T.AMF_OK
    assert conv.UpdateResponseTD(ir, L, None, T.TAN_CONVOLUTION_OPERATION_FLAG_BLOCK_UNTIL_READY) == T.AMF_OK
    nb = 80
    x = T.signals.input_bank(2, nb * B, base_seed=300)
    xd = torch.from_numpy(x).cuda()
    yd = torch.zeros_like(xd)
    for b in range(nb):
        rc, done = conv.Process(xd[:, b * B:], yd[:, b * B:], B)
        assert rc == T.AMF_OK and done == B
    conv.Synchronize()
    y = yd.cpu().numpy()
    conv.Terminate()
    _, ol, orr = O.room_response(room, src, lst, 48000, L)
    s = 2  # the first IR fades in at block 2
    for c, h in enumerate((ol, orr)):
        ref = fft_conv_f64(x[c], h)
        assert np.abs(y[c, (s + 1) * B:] - ref[(s + 1) * B:]).max() <= 1e-5
        assert snr_db(ref[(s + 1) * B:], y[c, (s + 1) * B:]) >= 120.0


def test_errors(gpu_ctx):
    lst = R.make_listener((1.0, 1.0, 1.0))
    src = R.MonoSource(1.0, 1.0, 1.0)
    a = np.zeros(1024, np.float32)
    b = np.zeros(1024, np.float32)
    assert R.generateRoomResponse(gpu_ctx, R.make_room(3, 3, 3, 1.5), src, lst, 48000, 1024, a, b) == 4
    assert R.generateRoomResponse(gpu_ctx, R.make_room(0.001, 3, 3), src, lst, 48000, 1024, a, b) == 4
    assert R.generateRoomResponse(gpu_ctx, R.make_room(0.01, 0.01, 0.01), src, lst, 48000, 1024, a, b) == 5
    lib = gpu_ctx._lib
    assert lib.tanx_room_response_device(gpu_ctx.handle, 0, None, None, None, 48000, 1024, 0, 0, None, 1024) != 0
