"""Copies the reference's convolution test inputs into tests/golden.

Run in a checkout that has the reference at /root/reference (data files
only; the reference is not executed).  TALibTestConvolution's own inputs
(tan/tests/src/TALibTestConvolution/): the two stereo responses
testresponse.wav / testresponse2.wav in full, and the first 24,000 frames
(0.5 s) of the stereo pulse train pulsetrainS.wav.
"""
import os
import shutil
import wave

SRC = "/root/reference/tan/tests/src/TALibTestConvolution"
DST = os.path.dirname(os.path.abspath(__file__))


def main():
    for name in ("testresponse.wav", "testresponse2.wav"):
        shutil.copyfile(os.path.join(SRC, name), os.path.join(DST, name))
    with wave.open(os.path.join(SRC, "pulsetrainS.wav")) as w:
        params = w.getparams()
        frames = w.readframes(24000)
    with wave.open(os.path.join(DST, "pulsetrainS_0p5s.wav"), "wb") as o:
        o.setnchannels(params.nchannels)
        o.setsampwidth(params.sampwidth)
        o.setframerate(params.framerate)
        o.writeframes(frames)


if __name__ == "__main__":
    main()
