"""C ABI: the library loads, exports every declared symbol, validates its
arguments on the host (no GPU needed for any of these calls)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from supra_amd import _native as N
from supra_amd.geometry import BeamformerConfig, RxGeometry

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(supra_\w+)\s*\(", text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_every_declared_symbol_is_exported(native):
    declared = declared_functions()
    assert len(declared) >= 10
    missing = [n for n in declared if not hasattr(native, n)]
    assert not missing, f"declared but not exported: {missing}"
    # and the binding table covers them
    assert set(declared) <= set(N.exported_symbols())


def test_version_and_error_strings(native):
    assert b"gfx950" in native.supra_version()
    assert native.supra_last_error() is not None


def test_window_lut_rejects_bad_type(native):
    out = np.zeros(64, dtype=np.float32)
    assert native.supra_window_lut(7, 0.0, N.ptr(out, C.c_float)) == -1
    assert b"window" in native.supra_last_error()
    assert native.supra_window_lut(1, 0.0, None) == -1


def test_create_rejects_bad_geometry(native):
    g = N.RxGeometryC()
    ctx = C.c_void_p()
    assert native.supra_bf_create(C.byref(g), C.byref(ctx)) == -1
    assert not ctx.value


def test_das_rejects_null_context(native):
    assert native.supra_das_run(None, 0, None, 1, 1, 1, 1, 40e6, 1.0, 1.54e6, 1, 0.0, 0, 0, None, 2, None) == -1


def test_geometry_rejects_bad_config(native):
    cfg = BeamformerConfig(num_scanlines=(1, 1)).to_c()
    n, ne = C.c_uint32(), C.c_uint32()
    lay = (C.c_uint32 * 2)()
    assert native.supra_geometry_sizes(C.byref(cfg), C.byref(n), lay, C.byref(ne)) == -1
    cfg = BeamformerConfig(max_aperture=(200, 1)).to_c()
    assert native.supra_geometry_sizes(C.byref(cfg), C.byref(n), lay, C.byref(ne)) == -1
    cfg = BeamformerConfig(scan_type="planewave", transducer_type="planar", element_layout=(8, 8),
                           max_aperture=(8, 1), tx_max_aperture=(8, 1)).to_c()
    assert native.supra_geometry_sizes(C.byref(cfg), C.byref(n), lay, C.byref(ne)) == -2


def test_struct_layouts_match_header():
    assert C.sizeof(N.TxParams) == 24
    assert C.sizeof(N.ScanlineRx) == 192
    assert C.sizeof(N.RxGeometryC) == 88


def test_library_window_matches_oracle():
    from oracle import oracle as O
    from supra_amd.beamformer import window_lut
    for wt, p in [(0, 0.0), (1, 0.0), (1, 0.3), (2, 0.0), (3, 0.5)]:
        assert np.array_equal(window_lut(wt, p), O.window_lut(wt, p, 32))


def test_scanconv_rejects_bad_arguments(native):
    from supra_amd.geometry import SCANLINE_DTYPE
    s = np.zeros(4, SCANLINE_DTYPE)
    sc = C.c_void_p()
    for layout, depth, ns, flags in [((1, 1), 10.0, 100, 0), ((1, 4), 10.0, 100, 0), ((4, 1), 0.0, 100, 0),
                                     ((4, 1), 10.0, 1, 0), ((4, 1), 10.0, 100, 2), ((0, 1), 10.0, 100, 0)]:
        lay = (C.c_uint32 * 2)(*layout)
        assert native.supra_scanconv_create(s.ctypes.data, lay, depth, ns, 0.0, flags, C.byref(sc)) == -1
        assert not sc.value
        assert b"scanconv" in native.supra_last_error()
    assert native.supra_scanconv_run(None, None, 2, 1, None, 2, None) == -1
    assert native.supra_scanconv_tables(None, None, None, None) == -1
    assert native.supra_scanconv_geometry(None, None, None, None, None) == -1
