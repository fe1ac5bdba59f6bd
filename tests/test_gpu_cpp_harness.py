"""The reference's convolution driver restated in C++ against
include/TrueAudioNext.h (tests/cpp/talib_harness.cpp), run on the driver's
own WAV inputs (tests/golden), read and written through the library's WAV
I/O; it checks every block against the oracle."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = os.path.join(ROOT, "tests", "cpp", "build", "talib_harness")


@pytest.mark.gpu
def test_talib_convolution_driver_through_cpp_interface(tmp_path):
    assert os.path.exists(HARNESS), "build() compiles the harness"
    print("running", HARNESS, flush=True)
    p = subprocess.run([HARNESS, os.path.join(ROOT, "tests", "golden"), str(tmp_path / "out.wav")],
                       capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr, flush=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "PASS" in p.stdout
