"""The oracle's room-response restatement (oracle/tan_oracle_room.c), pinned
on the CPU without the reference:

* the head filters against numpy's inverse FFT of the filtered spectrum,
  their DC gain and complementarity;
* the ear set-up against an FP64 rotation of the ear offsets;
* every image against a textbook image-source model built from reflection
  geometry: the image in cell c of an axis is the source reflected across
  the cell boundaries between cell 0 and cell c, and its reflections off
  each wall are the crossings of planes at even (near wall) and odd (far
  wall) multiples of the room size;
* the response assembled from the per-image contributions in 2^26 fixed
  point, bit for bit, and the direct path alone when every wall absorbs.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as O


class Room(ctypes.Structure):
    _fields_ = [(k, ctypes.c_float) for k in ("width", "height", "length", "damp_left", "damp_right", "damp_top",
                                                "damp_bottom", "damp_front", "damp_back")]


class Source(ctypes.Structure):
    _fields_ = [(k, ctypes.c_float) for k in ("x", "y", "z")]


class Listener(ctypes.Structure):
    _fields_ = ([(k, ctypes.c_float) for k in ("head_x", "head_y", "head_z", "ear_spacing", "yaw", "pitch",
                                               "roll")]
                + [("filter_length", ctypes.c_int32), ("low_pass", ctypes.c_float * 64),
                   ("high_pass", ctypes.c_float * 64)])


def listener(head, spacing=0.16, yaw=0.0, pitch=0.0, roll=0.0):
    lst = Listener(*head, spacing, yaw, pitch, roll)
    assert O.room_head_model(spacing, lst) == 0
    return lst


def test_head_model_against_numpy_ifft():
    for spacing in (0.12, 0.16, 0.2):
        lst = Listener()
        assert O.room_head_model(spacing, lst) == 0
        assert lst.filter_length == 64
        fc = np.float32(343.0 / (1.10 * np.float32(spacing)))
        j = np.arange(64)
        f = (j // 2).astype(np.float32) / (np.float32(64) * np.float32(1.0 / np.float32(48000)))
        d = (fc / (fc + f)).astype(np.float64)
        lp = np.fft.ifft(d).real  # numpy's inverse FFT includes the 1/64
        got_lp = np.array(lst.low_pass[:], np.float64)
        got_hp = np.array(lst.high_pass[:], np.float64)
        np.testing.assert_allclose(got_lp, lp, atol=2e-7, rtol=0)
        assert abs(got_lp.sum() - 1.0) < 1e-6  # DC gain d_0 = 1
        delta = np.zeros(64)
        delta[0] = 1.0
        np.testing.assert_allclose(got_lp + got_hp, delta, atol=1e-7, rtol=0)


def _rotation(yaw, pitch, roll):
    y, p, r = np.radians([yaw, pitch, roll])
    rz = np.array([[np.cos(r), -np.sin(r), 0], [np.sin(r), np.cos(r), 0], [0, 0, 1]])
    rx = np.array([[1, 0, 0], [0, np.cos(p), -np.sin(p)], [0, np.sin(p), np.cos(p)]])
    ry = np.array([[np.cos(y), 0, np.sin(y)], [0, 1, 0], [-np.sin(y), 0, np.cos(y)]])
    return rz @ rx @ ry  # yaw, then pitch, then roll


def _image_axis(c, size, s):
    """Position and (near, far) wall reflections of the image in cell c."""
    p, near, far = float(s), 0, 0
    step = 1 if c > 0 else -1
    for k in range(0, c, step):
        plane = (k + 1) * size if step > 0 else k * size
        p = 2 * plane - p
        m = round(plane / size)
        if m % 2 == 0:
            near += 1
        else:
            far += 1
    return p, near, far


CASES = [
    # room (w, h, l, damps L R T B F Bk), source, head, yaw/pitch/roll, length, flags, bounces
    ((5.0, 3.0, 4.0, 0.6, 0.7, 0.5, 0.8, 0.65, 0.55), (1.5, 1.2, 2.0), (3.0, 1.6, 2.5), (0, 0, 0), 4096, 0, 0),
    ((4.3, 2.7, 6.1, 0.9, 0.9, 0.3, 0.4, 0.8, 0.85), (0.5, 2.0, 5.0), (2.1, 1.7, 1.0), (35, -10, 5), 3000, 0, 0),
    ((3.0, 3.0, 3.0, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5), (9.0, -1.0, 1.0), (1.0, 1.5, 2.0), (90, 0, 0), 2048, 0x40, 2),
    ((6.0, 2.5, 4.5, 1.0, 0.2, 1.0, 0.0, 0.7, 0.3), (2.0, 1.0, 1.0), (4.0, 1.2, 3.5), (-120, 20, 0), 2500, 0x80, 0),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_images_against_textbook_model(case):
    dims, src, head, ang, n_len, flags, bounces = CASES[case]
    room, source = Room(*dims), Source(*src)
    lst = listener(head, 0.16, *ang)
    rc, ears, images = O.room_ear_params(room, source, lst, 48000, n_len, flags, bounces)
    assert rc == 0
    size = np.array(dims[:3], np.float64)
    s = np.clip(np.array(src, np.float64), 0, size)  # the source is kept inside the room
    R = _rotation(*ang)
    off = R @ np.array([[0.08, -0.08], [0, 0], [0, 0]])
    vec = R @ np.array([[0.08, -0.08], [0, 0], [-0.08, -0.08]])
    damp = dict(zip(("left", "right", "top", "bottom", "front", "back"), dims[3:]))
    near_w = [damp["left"], damp["bottom"], damp["back"]]
    far_w = [damp["right"], damp["top"], damp["front"]]
    n_img = 0
    for ch in range(2):
        E = ears[ch]
        ear = np.array(head, np.float64) + off[:, ch]
        np.testing.assert_allclose(np.array(E.head[:]), ear, atol=2e-6)
        np.testing.assert_allclose(np.array(E.ev[:]), vec[:, ch], atol=2e-7)
        ns = [int(size[k] / 340.0 * 48000) for k in range(3)]
        n = [1 + n_len // v for v in ns]
        if flags & 0x40:
            n = [min(v, bounces) for v in n]
        assert list(E.n) == n and images == np.prod([2 * v + 1 for v in n])
        assert E.hrtf_len == 2 * sum(ns)
        for cz in range(-n[2], n[2] + 1):
            for cy in range(-n[1], n[1] + 1):
                for cx in range(-n[0], n[0] + 1):
                    kind, idx, taps, scale, hf = O.room_image(E, cx, cy, cz)
                    if flags & 0x80 and cx == cy == cz == 0:
                        assert kind == 0
                        continue
                    pos, amp = np.zeros(3), 1.0
                    for k, c in enumerate((cx, cy, cz)):
                        pos[k], nr, fr = _image_axis(c, size[k], s[k])
                        amp *= near_w[k] ** nr * far_w[k] ** fr
                    d_vec = pos - ear
                    d = np.linalg.norm(d_vec)
                    t = d / 340.0 * 48000
                    want = 1 + int(np.floor(t))
                    if want >= n_len:
                        assert kind == 0 or abs(t - round(t)) < 1e-4
                        continue
                    n_img += 1
                    if abs(t - round(t)) < 1e-4:
                        assert abs(idx - want) <= 1
                    else:
                        assert idx == want, (cx, cy, cz)
                    dmin = 2 * 0.16
                    dr = 2.0 if d <= dmin else 2.0 * dmin / d
                    np.testing.assert_allclose(scale, 2.0 ** 26 * amp * dr, rtol=2e-6, atol=1e-3)
                    cosa = vec[:, ch] @ d_vec / (np.linalg.norm(vec[:, ch]) * d)
                    assert abs(hf - (cosa + 1) / 2) < 2e-6
                    assert kind == (2 if idx < E.hrtf_len else 1)
                    if kind == 2:
                        assert taps == min(64, n_len - idx)
    assert n_img > 100


@pytest.mark.parametrize("case", range(len(CASES)))
def test_response_is_the_fixed_point_sum_of_the_images(case):
    dims, src, head, ang, n_len, flags, bounces = CASES[case]
    room, source = Room(*dims), Source(*src)
    lst = listener(head, 0.16, *ang)
    rc, left, right = O.room_response(room, source, lst, 48000, n_len, flags, bounces)
    assert rc == 0
    _, ears, _ = O.room_ear_params(room, source, lst, 48000, n_len, flags, bounces)
    for ch, got in enumerate((left, right)):
        E = ears[ch]
        lp = np.array(E.lp[:], np.float32)
        hp = np.array(E.hp[:], np.float32)
        acc = np.zeros(n_len, np.int64)
        for cz in range(-E.n[2], E.n[2] + 1):
            for cy in range(-E.n[1], E.n[1] + 1):
                for cx in range(-E.n[0], E.n[0] + 1):
                    kind, idx, taps, scale, hf = O.room_image(E, cx, cy, cz)
                    if kind == 1:
                        acc[idx] += int(np.float32(scale))
                    elif kind == 2:
                        v = np.float32(scale) * (lp[:taps] + np.float32(hf) * hp[:taps])
                        acc[idx:idx + taps] += np.trunc(v).astype(np.int64)
        wrapped = ((acc + 2 ** 31) % 2 ** 32 - 2 ** 31).astype(np.float32)
        np.testing.assert_array_equal(got, wrapped / np.float32(2.0 ** 26))


def test_absorbing_walls_leave_the_direct_path():
    room = Room(5.0, 3.0, 4.0, *([0.0] * 6))
    head = (3.0, 1.6, 2.5)
    lst = listener(head)
    src = (1.5, 1.2, 2.0)
    rc, left, right = O.room_response(room, Source(*src), lst, 48000, 4096)
    assert rc == 0
    lp = np.array(lst.low_pass[:], np.float64)
    hp = np.array(lst.high_pass[:], np.float64)
    for ch, got in enumerate((left, right)):
        ear = np.array(head) + np.array([0.08 if ch == 0 else -0.08, 0, 0])
        v = np.array([0.08, 0, -0.08]) if ch == 0 else np.array([-0.08, 0, -0.08])
        dv = np.array(src) - ear
        d = np.linalg.norm(dv)
        idx = 1 + int(d / 340.0 * 48000)
        hf = (v @ dv / (np.linalg.norm(v) * d) + 1) / 2
        want = np.zeros(4096)
        want[idx:idx + 64] = 2.0 * 0.32 / d * (lp + hf * hp)
        np.testing.assert_allclose(got, want, atol=3e-7, rtol=1e-5)
        assert np.count_nonzero(got[:idx]) == 0 and np.count_nonzero(got[idx + 64:]) == 0


def test_invalid_arguments():
    lst = listener((1.0, 1.0, 1.0))
    src = Source(1.0, 1.0, 1.0)
    for room in (Room(0.0, 3, 3, *([0.5] * 6)), Room(3, 3, 3, 1.5, *([0.5] * 5)),
                 Room(0.001, 3, 3, *([0.5] * 6))):
        rc, _, _ = O.room_response(room, src, lst, 48000, 1024)
        assert rc == 4
    rc, _, _ = O.room_ear_params(Room(0.01, 0.01, 0.01, *([0.5] * 6)), src, lst, 48000, 1 << 20)
    assert rc == 5  # more than 2^31 - 1 images
