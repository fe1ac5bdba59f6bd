"""WAV I/O of the library (tanx_wav_*, tan/common/wav.cpp semantics) on the
CPU: the reference's own test inputs decoded against Python's wave module,
hand-built 8-bit, 32-bit float and multi-chunk files, the writers' bytes
against a hand-packed canonical header, and the error paths."""
import os
import struct
import wave

import numpy as np
import pytest

from tan_amd import wav

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _wave_module_samples(path):
    with wave.open(path, "rb") as w:
        ch, width, n = w.getnchannels(), w.getsampwidth(), w.getnframes()
        raw = w.readframes(n)
    assert width == 2
    return ch, np.frombuffer(raw, "<i2").reshape(n, ch).T


def _riff(chunks):
    body = b"WAVE" + b"".join(cid + struct.pack("<I", len(data)) + data + (b"\0" if len(data) & 1 else b"")
                              for cid, data in chunks)
    return b"RIFF" + struct.pack("<I", len(body)) + body


def _fmt(tag, ch, rate, bits, extra=b""):
    return struct.pack("<HHIIHH", tag, ch, rate, rate * ch * bits // 8, ch * bits // 8, bits) + extra


@pytest.mark.parametrize("name", ["testresponse.wav", "testresponse2.wav", "pulsetrainS_0p5s.wav"])
def test_reference_inputs_decode_like_the_reference_reader(name):
    path = os.path.join(GOLDEN, name)
    info, x = wav.ReadWaveFile(path)
    ch, s = _wave_module_samples(path)
    with wave.open(path, "rb") as w:
        rate = w.getframerate()
    assert info == {"sample_rate": rate, "bits_per_sample": 16, "channels": ch, "samples": s.shape[1],
                    "format_tag": 1}
    # wav.cpp:214-223: (float)s / 32768.0f
    np.testing.assert_array_equal(x, s.astype(np.float32) / np.float32(32768.0))
    rinfo, raw = wav.ReadWaveRaw(path)
    assert rinfo == info
    assert raw == s.T.astype("<i2").tobytes()


def test_8bit_32bit_and_extra_chunks(tmp_path):
    rng = np.random.default_rng(5)
    u8 = rng.integers(0, 256, size=(3, 50)).astype(np.uint8)
    p8 = tmp_path / "u8.wav"
    # a LIST chunk before "fmt ", an 18-byte fmt, a "fact" chunk before "data"
    p8.write_bytes(_riff([(b"LIST", b"INFOxyz"), (b"fmt ", _fmt(1, 3, 22050, 8, b"\0\0")),
                          (b"fact", b"\1\2\3\4"), (b"data", u8.T.tobytes())]))
    info, x = wav.ReadWaveFile(p8)
    assert (info["channels"], info["bits_per_sample"], info["samples"], info["sample_rate"]) == (3, 8, 50, 22050)
    # wav.cpp:205-213: (float)(u - 127) / 256.0f
    np.testing.assert_array_equal(x, (u8.astype(np.int32) - 127).astype(np.float32) / np.float32(256.0))

    f32 = rng.standard_normal((2, 40)).astype(np.float32)
    pf = tmp_path / "f32.wav"
    pf.write_bytes(_riff([(b"fmt ", _fmt(3, 2, 48000, 32)), (b"data", f32.T.astype("<f4").tobytes())]))
    info, x = wav.ReadWaveFile(pf)
    assert info["format_tag"] == 3 and info["samples"] == 40
    np.testing.assert_array_equal(x, f32)


def test_truncated_data_reads_as_zero(tmp_path):
    s = np.arange(1, 21, dtype="<i2")
    p = tmp_path / "t.wav"
    blob = _riff([(b"fmt ", _fmt(1, 1, 8000, 16)), (b"data", s.tobytes())])
    p.write_bytes(blob[:-10])  # the last 5 samples are missing
    info, x = wav.ReadWaveFile(p)
    assert info["samples"] == 20
    np.testing.assert_array_equal(x[0, :15], s[:15].astype(np.float32) / 32768.0)
    np.testing.assert_array_equal(x[0, 15:], 0.0)


def test_writer_bytes_match_the_canonical_header_and_conversions(tmp_path):
    x = np.array([[0.0, 0.5, -0.5, 1.0, -1.0, 1.7, -3.0, 0.99999, -0.123456, 3.1e-5],
                  [0.25, -0.25, 0.75, -0.75, 0.1, -0.1, 0.0, 2.0, -2.0, 0.333]], np.float32)
    for bits, width, tag in ((16, 2, 1), (8, 1, 1), (32, 4, 3)):
        p = tmp_path / f"w{bits}.wav"
        wav.WriteWaveFileF(p, 44100, bits, x)
        blob = p.read_bytes()
        n = x.shape[1] * x.shape[0] * width
        hdr = (b"RIFF" + struct.pack("<I", n + 36) + b"WAVEfmt " + struct.pack("<I", 16)
               + struct.pack("<HHIIHH", tag, 2, 44100, 2 * 44100 * width, 2 * width, bits)
               + b"data" + struct.pack("<I", n))
        assert blob[:44] == hdr, bits
        xc = np.clip(x.astype(np.float64), -1.0, 1.0).T.ravel()
        if bits == 16:  # wav.cpp:284-296: (short)(value * 32767.0)
            want = np.trunc(xc * 32767.0).astype("<i2").tobytes()
        elif bits == 8:  # wav.cpp:271-283: (char)(value * 127.0)
            want = np.trunc(xc * 127.0).astype(np.int8).tobytes()
        else:
            want = x.T.astype("<f4").tobytes()
        assert blob[44:] == want, bits
    # a 16-bit file read back: trunc(v * 32767) / 32768
    info, y = wav.ReadWaveFile(tmp_path / "w16.wav")
    np.testing.assert_array_equal(
        y, (np.trunc(np.clip(x.astype(np.float64), -1, 1) * 32767.0) / 32768.0).astype(np.float32))
    # Python's wave module reads what the writer wrote
    with wave.open(str(tmp_path / "w16.wav"), "rb") as w:
        assert (w.getnchannels(), w.getsampwidth(), w.getframerate(), w.getnframes()) == (2, 2, 44100, 10)


def test_short_writer(tmp_path):
    s = np.array([1, -2, 300, -32768, 32767, 0], np.int16)
    p = tmp_path / "s.wav"
    wav.WriteWaveFileS(p, 16000, 2, s)
    blob = p.read_bytes()
    assert blob[44:] == s.astype("<i2").tobytes()
    assert wav.ReadWaveInfo(p) == {"sample_rate": 16000, "bits_per_sample": 16, "channels": 2, "samples": 3,
                                   "format_tag": 1}
    with pytest.raises(wav.WavError) as e:
        wav.WriteWaveFileS(p, 16000, 2, s, bits_per_sample=8)
    assert e.value.rc == 9  # AMF_NOT_IMPLEMENTED (the reference throws)


def test_errors(tmp_path):
    with pytest.raises(wav.WavError) as e:
        wav.ReadWaveFile(tmp_path / "missing.wav")
    assert e.value.rc == 16  # AMF_FILE_NOT_OPEN
    bad = tmp_path / "bad.wav"
    bad.write_bytes(b"RIFX" + b"\0" * 40)
    with pytest.raises(wav.WavError) as e:
        wav.ReadWaveFile(bad)
    assert e.value.rc == 14  # AMF_INVALID_FORMAT
    nodata = tmp_path / "nodata.wav"
    nodata.write_bytes(_riff([(b"fmt ", _fmt(1, 1, 8000, 16))]))
    with pytest.raises(wav.WavError) as e:
        wav.ReadWaveFile(nodata)
    assert e.value.rc == 14
    b24 = tmp_path / "b24.wav"
    b24.write_bytes(_riff([(b"fmt ", _fmt(1, 1, 8000, 24)), (b"data", b"\0" * 9)]))
    with pytest.raises(wav.WavError) as e:
        wav.ReadWaveFile(b24)
    assert e.value.rc == 14
