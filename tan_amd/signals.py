"""Deterministic synthetic signals for the convolution workloads.

IR model: Gaussian noise times exp(-t / 0.25 s) at 48 kHz, normalised to
unit L2; input: white noise uniform in [-1, 1].  Every array is generated in
float64 from numpy's PCG64 seeded by (base seed, stream kind, channel,
update index) and rounded once to float32, so the GPU path and the CPU
oracle see identical float32 inputs.
"""
import numpy as np

SAMPLE_RATE = 48000
DECAY_SECONDS = 0.25
KIND_IR = 1
KIND_INPUT = 2


def _rng(base, kind, channel, update=0):
    return np.random.Generator(np.random.PCG64([int(base), kind, int(channel), int(update)]))


def decaying_noise_ir(length, channel, update=0, base_seed=1234, tau_s=DECAY_SECONDS):
    g = _rng(base_seed, KIND_IR, channel, update).standard_normal(length)
    h = g * np.exp(-np.arange(length) / (tau_s * SAMPLE_RATE))
    h /= np.sqrt(np.sum(h * h))
    return h.astype(np.float32)


def white_noise(length, channel, base_seed=1234):
    return _rng(base_seed, KIND_INPUT, channel).uniform(-1.0, 1.0, length).astype(np.float32)


def ir_bank(channels, length, update=0, base_seed=1234):
    return np.stack([decaying_noise_ir(length, c, update, base_seed) for c in range(channels)])


def input_bank(channels, length, base_seed=1234):
    return np.stack([white_noise(length, c, base_seed) for c in range(channels)])
