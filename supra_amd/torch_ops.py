"""The libtorch autograd operator of the differentiable 2D delay-and-sum.

``libsupra_torch.so`` (include/supra_bf_torch.hpp, csrc/torch_op.cpp)
registers ``torch.ops.supra_bf.das_c``: the torch::autograd::Function over
supra_das_forward_c / supra_das_adjoint that C++ libtorch programs use
directly.  ``das_c(engine, rf, c, ...)`` calls it from Python; it returns
the same image and gradients as ``DifferentiableDAS`` (the same C ABI).
"""
from __future__ import annotations

from pathlib import Path

import torch

from . import _native as N
from .beamformer import RxBeamformer

LIB = Path(__file__).resolve().parent / "libsupra_torch.so"
_loaded = False


def load():
    """Load libsupra_torch.so once; raise if it has not been built."""
    global _loaded
    if not _loaded:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is missing: build it with `python -m supra_amd.build`")
        torch.ops.load_library(str(LIB))
        _loaded = True
    return torch.ops.supra_bf


def das_c(engine: RxBeamformer, rf, c_m_per_s, fs_hz: float, f_number: float = 1.0,
          window_type: int = N.WINDOW_HANN, window_parameter: float = 0.0, interpolate_transmits: bool = False,
          additional_offset: int = 0):
    """Image [frame][depth][scanline] of rf [frame][tx][ch][s] (float32,
    device) at one speed of sound per frame (m/s), differentiable in rf and c."""
    ops = load()
    g = engine.geometry
    squeeze = rf.dim() == 3
    x = rf.unsqueeze(0) if squeeze else rf
    c = torch.as_tensor(c_m_per_s, device=rf.device)
    if c.dim() == 0:
        c = c.expand(x.shape[0])
    img = ops.das_c(x, c, int(engine._ctx.value), float(fs_hz), float(f_number), int(window_type),
                    float(window_parameter), bool(interpolate_transmits), int(additional_offset), g.num_depths,
                    g.num_rx)
    return img[0] if squeeze else img
