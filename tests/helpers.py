"""Shared comparison helpers for the parity tests."""
import numpy as np


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def rel_err(a, b, floor=1e-3):
    """Per-pixel relative error of RGB radiance (max over channels), with an absolute floor so that
    near-black pixels compare absolutely."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return (np.abs(a - b) / np.maximum(np.abs(b), floor)).max(axis=-1)


def rel_mse(x, ref, eps=1e-2):
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + eps)))


# Divergent-path fractions measured by the GPU parity tests, written out at the end of the session
# (tests/conftest.py) so every run leaves a summary of the measured fractions next to their bars.
DIVERGENCE = []


def record_divergence(name, frac, bar, pixels):
    DIVERGENCE.append((name, float(frac), float(bar), int(pixels)))
