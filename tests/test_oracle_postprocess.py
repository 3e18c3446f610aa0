"""Pins the host oracle's envelope detection (HilbertFirEnvelope.cu) and log
compression (LogCompressor.cu): the Hilbert FIR against its closed form, the
envelope of a pure tone against the analytic signal, log compression against
its closed form and end points."""
import math

import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("n", [65, 31, 257, 1])
def test_hilbert_filter_closed_form(n):
    h = O.hilbert_filter(n, precision=64)
    c = (n - 1) // 2
    expect = np.zeros(n)
    for i in range(n):
        k = i - c
        if k % 2:
            expect[i] = 2.0 / (math.pi * k)
        expect[i] *= 0.54 - 0.46 * math.cos(2 * math.pi * i / (n - 1)) if n > 1 else float("nan")
    if n == 1:
        assert np.isnan(h[0])          # the reference's window divides by N - 1 = 0
        return
    assert np.allclose(h, expect, rtol=1e-15, atol=0)
    assert np.all(h[c::2] == 0) and np.all(h[c::-2] == 0)       # even offsets
    assert np.allclose(h[:c][::-1], -h[c + 1:], rtol=1e-15)      # antisymmetric
    h32 = O.hilbert_filter(n)
    assert np.isfinite(h32).all()
    assert np.abs(h32 - h).max() <= 1e-7


@pytest.mark.parametrize("omega", [0.9, 1.6, 2.3])
def test_envelope_of_pure_tone_is_its_amplitude(omega):
    """x[t] = A cos(w t + phi): the FIR Hilbert transformer returns A sin(...)
    within its passband ripple, so sqrt(x^2 + H{x}^2) = A away from the ends."""
    n_s, n_rx = 600, 5
    rng = np.random.default_rng(3)
    amp = rng.uniform(100.0, 1000.0, n_rx)
    phi = rng.uniform(0, 2 * np.pi, n_rx)
    t = np.arange(n_s)[:, None]
    x = (amp[None, :] * np.cos(omega * t + phi[None, :])).astype(np.float32)
    env = O.envelope(x, 65, precision=64)
    inner = env[40:-40]
    assert np.abs(inner / amp[None, :] - 1).max() < 2e-2
    e32 = O.envelope(x, 65)
    assert np.abs(e32[40:-40] - inner).max() <= 1e-4 * amp.max()


def test_envelope_constant_and_edges():
    """A constant c: the antisymmetric FIR cancels inside the image (|c|); at
    the ends the truncated window leaves the one-sided sum."""
    x = np.full((200, 3), 7.0, dtype=np.float32)
    env = O.envelope(x, 65, precision=64)
    assert np.allclose(env[32:-32], 7.0, rtol=1e-12)
    h = O.hilbert_filter(65, precision=64)
    acc0 = 7.0 * h[32:].sum()                # sample 0 sees taps 32..64
    assert abs(env[0, 0] - math.hypot(7.0, acc0)) < 1e-9


def test_logcompress_closed_form():
    rng = np.random.default_rng(5)
    v = rng.uniform(-5000, 5000, 1000).astype(np.float32)
    for db, gain, in_max in [(80.0, 1.0, 1024.0), (40.0, 0.5, 300.0), (120.0, 2.0, 50000.0)]:
        y = O.logcompress(v, db, gain, in_max, precision=64)
        dr = 10 ** (db / 20)
        expect = np.log10(np.abs(v.astype(np.float64)) * (dr / in_max) + 1) * gain * 255 / np.log10(dr + 1)
        assert np.allclose(y, expect, rtol=1e-12)
        ends = O.logcompress(np.array([0.0, in_max, -in_max], np.float32), db, gain, in_max, precision=64)
        assert ends[0] == 0 and np.allclose(ends[1:], gain * 255, rtol=1e-12)
        y32 = O.logcompress(v, db, gain, in_max)
        assert np.abs(y32 - y).max() <= 1e-5 * gain * 255


def test_logcompress_uint8_in_max_wraps_like_the_reference():
    """The functor takes inMax as the input element type: for uint8 input an
    inMax of 1024 becomes 0, so every non-zero input saturates."""
    v = np.array([0, 1, 200], dtype=np.uint8)
    y = O.logcompress(v, 80.0, 1.0, 1024.0)
    assert np.isnan(y[0]) and np.isinf(y[1]) and np.isinf(y[2])
    assert O.clamp_cast(y, np.uint8).tolist() == [0, 255, 255]
