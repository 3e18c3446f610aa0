"""Writes the small regression fixtures of tests/golden/ from the host oracle.

Run from the repository root: python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402

HERE = Path(__file__).resolve().parent


def main():
    luts = {}
    for wt, p in [(0, 0.0), (1, 0.0), (1, 0.25), (2, 0.0), (2, 0.5), (3, 0.3), (3, 1.5)]:
        luts[f"{wt}:{p}"] = [float(v) for v in O.window_lut(wt, p, 32)]
    (HERE / "window_luts.json").write_text(json.dumps(luts, indent=0))
    print("wrote", HERE / "window_luts.json")


if __name__ == "__main__":
    main()
