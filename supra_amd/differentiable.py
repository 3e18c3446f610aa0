"""Differentiable 2D delay-and-sum for PyTorch.

``DifferentiableDAS(geometry, fs_hz, ...)(rf, c)`` beamforms a batch of
frames, each with its own speed of sound, as an autograd-aware operation on
device tensors.  Gradients with respect to the RF and to the speed of sound
come from the hand-written adjoint in libsupra_bf.so (supra_das_forward_c /
supra_das_adjoint); the delays are recomputed in the backward pass, only the
per-pixel normalisation n/W is kept from the forward.

The reference links libtorch behind its optional SUPRA_TORCH switch
(CMakeLists.txt:32,74-77; src/SupraLib/CMakeLists.txt:342-356;
Processing/TorchInference.h:27-294) for inference only; this module is the
optional torch-facing layer of the beamformer, with the arguments of the
DelayAndSum path of RxBeamformerCuda::performRxBeamforming
(RxBeamformerCuda.cu:539-663) except that the speed of sound is a tensor
(m/s, one value per frame).
"""
from __future__ import annotations

import ctypes as C

from . import _native as N
from .beamformer import RxBeamformer, _stream_handle
from .geometry import RxGeometry

import torch


def _vp(t) -> C.c_void_p:
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


class DifferentiableDAS:
    """2D DelayAndSum with per-frame speed of sound and a hand-written adjoint."""

    def __init__(self, geometry: RxGeometry, fs_hz: float, f_number: float = 1.0,
                 window_type: int = N.WINDOW_HANN, window_parameter: float = 0.0,
                 interpolate_transmits: bool = False, additional_offset: int = 0):
        if geometry.is3d:
            raise ValueError("the differentiable beamformer takes 2D geometries")
        self.geometry = geometry
        self.engine = RxBeamformer(geometry)
        self.fs_hz = float(fs_hz)
        self.f_number = float(f_number)
        self.window_type = int(window_type)
        self.window_parameter = float(window_parameter)
        self.interpolate_transmits = bool(interpolate_transmits)
        self.additional_offset = int(additional_offset)

    @property
    def slots(self) -> int:
        return 4 if self.interpolate_transmits else 1

    def close(self) -> None:
        self.engine.close()

    @staticmethod
    def _frames(rf, c):
        if rf.dim() == 3:
            rf = rf.unsqueeze(0)
        if rf.dim() != 4 or rf.dtype != torch.float32 or not rf.is_cuda or not rf.is_contiguous():
            raise ValueError("rf must be a contiguous float32 device tensor [frame][tx][ch][s] or [tx][ch][s]")
        frames = int(rf.shape[0])
        c = torch.as_tensor(c, device=rf.device)
        if c.dim() == 0:
            c = c.expand(frames)
        if tuple(c.shape) != (frames,):
            raise ValueError(f"speed of sound must be a scalar or have shape ({frames},)")
        return rf, c, frames

    def _c_mm(self, c):
        return (c.detach().to(torch.float64) * 1000.0).contiguous()

    def forward_raw(self, rf, c_m_per_s, need_norm: bool = True, stream=None):
        """Image [frame][depth][scanline] and, if asked, the normalisation
        [frame][slot][scanline][depth] that the adjoint consumes."""
        rf, c, frames = self._frames(rf, c_m_per_s)
        g = self.geometry
        out = torch.empty((frames, g.num_depths, g.num_rx), dtype=torch.float32, device=rf.device)
        norm = torch.empty((frames, self.slots, g.num_rx, g.num_depths), dtype=torch.float32,
                           device=rf.device) if need_norm else None
        c_mm = self._c_mm(c)
        N.check(N.lib().supra_das_forward_c(
            self.engine._ctx, frames, _vp(rf), int(rf.shape[1]), int(rf.shape[2]), int(rf.shape[3]), self.fs_hz,
            self.f_number, _vp(c_mm), self.window_type, self.window_parameter, int(self.interpolate_transmits),
            self.additional_offset, _vp(out), _vp(norm), C.c_void_p(_stream_handle(stream))))
        return out, norm

    def adjoint_raw(self, rf, c_m_per_s, norm, grad_image, want_rf: bool = True, want_c: bool = True,
                    stream=None):
        """(dL/drf, dL/dc in 1/(m/s)) for dL/d(image) = grad_image."""
        rf, c, frames = self._frames(rf, c_m_per_s)
        g = self.geometry
        grad_image = grad_image.reshape(frames, g.num_depths, g.num_rx).to(torch.float32).contiguous()
        grad_rf = torch.empty_like(rf) if want_rf else None
        grad_c = torch.empty((frames,), dtype=torch.float64, device=rf.device) if want_c else None
        c_mm = self._c_mm(c)
        N.check(N.lib().supra_das_adjoint(
            self.engine._ctx, frames, _vp(rf), int(rf.shape[1]), int(rf.shape[2]), int(rf.shape[3]), self.fs_hz,
            self.f_number, _vp(c_mm), self.window_type, self.window_parameter, int(self.interpolate_transmits),
            self.additional_offset, _vp(norm), _vp(grad_image), _vp(grad_rf), _vp(grad_c),
            C.c_void_p(_stream_handle(stream))))
        if grad_c is not None:
            grad_c = grad_c * 1000.0          # 1/(mm/s) -> 1/(m/s)
        return grad_rf, grad_c

    def __call__(self, rf, c_m_per_s):
        """Autograd-aware beamforming: image [frame][depth][scanline] (or [depth][scanline] for 3D rf)."""
        squeeze = rf.dim() == 3
        rf4 = rf.unsqueeze(0) if squeeze else rf
        c = torch.as_tensor(c_m_per_s, device=rf.device)
        if c.dim() == 0:
            c = c.expand(rf4.shape[0])
        out = _DASFunction.apply(rf4, c, self)
        return out[0] if squeeze else out


class _DASFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rf, c, module):
        need = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
        image, norm = module.forward_raw(rf.contiguous(), c, need_norm=need)
        ctx.module = module
        ctx.save_for_backward(rf, c, norm)
        return image

    @staticmethod
    def backward(ctx, grad_image):
        rf, c, norm = ctx.saved_tensors
        want_rf, want_c = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        grad_rf, grad_c = ctx.module.adjoint_raw(rf.contiguous(), c, norm, grad_image, want_rf, want_c)
        if grad_c is not None:
            grad_c = grad_c.to(c.dtype)
        return grad_rf, grad_c, None
