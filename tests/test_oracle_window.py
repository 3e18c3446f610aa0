"""Pins the oracle's window table (WindowFunction.h:121-138, WindowFunction.cpp:18-47)
against closed forms evaluated independently in float64."""
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = Path(__file__).resolve().parent / "golden" / "window_luts.json"


def closed_form(wt, p):
    r = np.arange(64) / 63.0 * 2 - 1
    if wt == 0:
        v = np.ones(64)
    elif wt == 1:
        v = (1 - p) * (0.5 - 0.5 * np.cos(np.pi * (r + 1))) + p
    elif wt == 2:
        v = (1 - p) * (0.54 - 0.46 * np.cos(np.pi * (r + 1))) + p
    else:
        v = np.exp(-0.5 * (r / p) ** 2) / (p * np.sqrt(2 * np.pi))
    return v / v.max()


CASES = [(0, 0.0), (1, 0.0), (1, 0.25), (2, 0.0), (2, 0.5), (3, 0.3), (3, 1.5)]


@pytest.mark.parametrize("wt,p", CASES)
def test_window_matches_closed_form(wt, p):
    f32 = O.window_lut(wt, p, 32)
    f64 = O.window_lut(wt, p, 64)
    ref = closed_form(wt, np.float32(p).astype(np.float64))
    assert np.abs(f64 - ref).max() < 1e-12
    assert np.abs(f32 - ref).max() < 4e-7
    assert f32.max() == 1.0 and f64.max() == 1.0


def test_window_shapes():
    hann = O.window_lut(1, 0.0, 64)
    assert hann[0] == 0.0 and hann[63] == 0.0          # Hann ends at zero
    assert np.allclose(hann, hann[::-1], atol=1e-15)   # symmetric
    ham = O.window_lut(2, 0.0, 64)
    assert abs(ham[0] - 0.08 / (0.54 + 0.46 * np.cos(np.pi / 63))) < 1e-12   # peak at r = +-1/63
    assert np.all(O.window_lut(0, 0.0, 32) == 1.0)


def test_window_golden_tables():
    data = json.loads(GOLDEN.read_text())
    for key, vals in data.items():
        wt, p = key.split(":")
        got = O.window_lut(int(wt), float(p), 32)
        assert np.array_equal(got, np.array(vals, dtype=np.float32)), key
