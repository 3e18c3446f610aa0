"""The libtorch autograd operator supra_bf::das_c (include/supra_bf_torch.hpp):
it loads and registers its schema (host), and on the GPU it returns the image
and the RF / speed-of-sound gradients of DifferentiableDAS bit for bit (both
reach the same C ABI, one through a C++ torch::autograd::Function)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def test_operator_registers_its_schema():
    from supra_amd import torch_ops
    ops = torch_ops.load()
    schema = str(ops.das_c.default._schema)
    assert schema.startswith("supra_bf::das_c(Tensor rf, Tensor c, int handle")
    with pytest.raises(RuntimeError, match="null beamformer context"):
        ops.das_c(torch.zeros(1, 2, 3, 4), torch.zeros(1), 0, 40e6, 1.0, 1, 0.0, False, 0, 1, 1)


@pytest.mark.gpu
def test_operator_matches_differentiable_das(device):
    from supra_amd import _native as N
    from supra_amd import torch_ops
    from supra_amd import workloads as W
    from supra_amd.differentiable import DifferentiableDAS
    w = W.linear_2d(num_scatterers=60)
    g = w.geometry()
    frames = 3
    rf = torch.stack([torch.from_numpy(W.linear_2d(scatter_seed=40 + f, noise_seed=50 + f).rf_host(dtype=np.float32))
                      for f in range(frames)]).cuda()
    c0 = torch.tensor([1490.0, 1540.0, 1585.0], dtype=torch.float64, device="cuda")
    dd = DifferentiableDAS(g, w.fs_hz, 1.0, N.WINDOW_HANN)
    outs = []
    for use_op in (False, True):
        x = rf.clone().requires_grad_(True)
        c = c0.clone().requires_grad_(True)
        img = torch_ops.das_c(dd.engine, x, c, w.fs_hz, 1.0, N.WINDOW_HANN) if use_op else dd(x, c)
        weight = torch.linspace(-1.0, 1.0, img.numel(), device="cuda").reshape(img.shape)
        loss = (torch.where(img.abs() < 1e30, img, 0.0) * weight).sum()
        loss.backward()
        torch.cuda.synchronize()
        outs.append((img.detach().cpu().numpy(), x.grad.cpu().numpy(), c.grad.cpu().numpy()))
    (a_img, a_rf, a_c), (b_img, b_rf, b_c) = outs
    assert np.array_equal(a_img, b_img)
    assert np.array_equal(a_rf, b_rf)
    assert np.array_equal(a_c, b_c)
    assert np.abs(a_c).min() > 0
    dd.close()
