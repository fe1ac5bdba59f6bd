"""The oracle's output chain (mixer, converter) against closed forms: the
channel order of the sum, truncation toward zero after clamping, the clip
flag, gain, steps and offsets."""
import numpy as np

import oracle_lib as O


def test_mix_sums_in_channel_order():
    # (1e8 + 1) - 1e8 is 0 in single precision summed in order, 1 otherwise
    a = np.array([[1e8], [1.0], [-1e8]], np.float32)
    assert O.mix(a)[0] == 0.0
    rng = np.random.default_rng(1)
    x = rng.standard_normal((7, 300)).astype(np.float32)
    ref = x[0].copy()
    for c in range(1, 7):
        ref = (ref + x[c]).astype(np.float32)
    np.testing.assert_array_equal(O.mix(x), ref)
    np.testing.assert_array_equal(O.mix(x[:1]), x[0])


def test_float_to_short_truncates_and_clamps():
    x = np.array([0.5, -0.5, 1.0, -1.0, 1.5, -1.1, 0.0, 3.0e-5, -3.0e-5], np.float32)
    y = np.zeros(x.size, np.int16)
    clipped = O.convert(x, 1, 0, y, 1, 0, x.size, 1.0)
    # 16383.5 -> 16383 (toward zero; nearest-even would give 16384)
    assert list(y) == [16383, -16383, 32767, -32767, 32767, -32768, 0, 0, 0]
    assert clipped
    y2 = np.zeros(4, np.int16)
    assert not O.convert(x[:4], 1, 0, y2, 1, 0, 4, 1.0)
    # gain scales before clamping: 0.5 * 32767 * 2 = 32767
    y3 = np.zeros(1, np.int16)
    assert not O.convert(np.array([0.5], np.float32), 1, 0, y3, 1, 0, 1, 2.0) and y3[0] == 32767


def test_short_to_float_steps_offsets_and_copies():
    s = np.array([32767, -32768, 0, 16384], np.int16)
    f = np.zeros(4, np.float32)
    assert not O.convert(s, 1, 0, f, 1, 0, 4, 1.0)
    sc = np.float32(1.0) / np.float32(32767.0)
    np.testing.assert_array_equal(f, (s.astype(np.float32) * sc).astype(np.float32))
    assert f[0] == 1.0
    # interleave two mono channels into stereo: steps 2, offsets 0 / 1
    left = np.array([0.1, 0.2, 0.3], np.float32)
    right = np.array([-0.1, -0.2, -0.3], np.float32)
    st = np.full(6, 7, np.int16)
    O.convert(left, 1, 0, st, 2, 0, 3, 1.0)
    O.convert(right, 1, 0, st, 2, 1, 3, 1.0)
    assert list(st) == [3276, -3276, 6553, -6553, 9830, -9830]
    # same-type conversions copy, gain ignored
    g = np.zeros(3, np.float32)
    O.convert(left, 1, 0, g, 1, 0, 3, 5.0)
    np.testing.assert_array_equal(g, left)
