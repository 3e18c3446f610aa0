"""The C++ engine shim (include/supra_bf.hpp): its test program, built with
the library, converts SUPRA-shaped geometry to the C ABI field for field
(host) and beamforms bit-identically to supra_das_run (GPU)."""
import subprocess
from pathlib import Path

import pytest

PROG = Path(__file__).resolve().parent.parent / "supra_amd" / "shim_test"


def run(mode):
    if not PROG.exists():
        pytest.fail(f"{PROG} is missing: build the library (python -m supra_amd.build)")
    res = subprocess.run([str(PROG), mode], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    return res


def test_shim_converts_supra_geometry_on_the_host():
    res = run("host")
    assert res.returncode == 0, res.stdout + res.stderr
    assert "0 failures" in res.stdout


@pytest.mark.gpu
def test_shim_beamforms_like_the_c_abi():
    res = run("gpu")
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.count("shim == supra_das_run") == 2
