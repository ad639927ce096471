"""The closure oracle pinned on the CPU: its convolution against
torch.nn.functional.conv3d, its backward passes against autograd, its
network gradient against finite differences, and the BF16 rounding against
torch's own bfloat16 conversion."""

import numpy as np
import pytest
import torch

from oracle import closure_oracle as co


def _torch_weight(wf):
    """wf[t, co, ci] -> torch conv3d weight [co, ci, 3, 3, 3] (cross-correlation)."""
    w = np.asarray(wf).reshape(3, 3, 3, co.C, co.C)  # [di, dj, dk, co, ci]
    return torch.from_numpy(np.ascontiguousarray(w.transpose(3, 4, 0, 1, 2)))


def _torch_conv(x, wf, bias):
    xt = torch.from_numpy(np.ascontiguousarray(np.moveaxis(x, -1, 0)))[None]
    out = torch.nn.functional.conv3d(xt, _torch_weight(wf), torch.from_numpy(bias), padding=1)
    return np.moveaxis(out[0].numpy(), 0, -1)


@pytest.mark.parametrize("shape", [(3, 4, 5), (6, 2, 16), (1, 1, 7)])
def test_conv3d_matches_torch(shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.normal(size=shape + (co.C,))
    wf = rng.normal(size=(co.TAPS, co.C, co.C))
    b = rng.normal(size=co.C)
    np.testing.assert_allclose(co.conv3d(x, wf, b), _torch_conv(x, wf, b), rtol=1e-10, atol=1e-9)


def test_backward_passes_match_autograd():
    rng = np.random.default_rng(5)
    shape = (4, 3, 6)
    x = rng.normal(size=shape + (co.C,))
    wf = rng.normal(size=(co.TAPS, co.C, co.C))
    b = rng.normal(size=co.C)
    dout = rng.normal(size=shape + (co.C,))
    xt = torch.tensor(np.moveaxis(x, -1, 0)[None], requires_grad=True)
    wt = _torch_weight(wf).clone().requires_grad_(True)
    bt = torch.tensor(b, requires_grad=True)
    out = torch.nn.functional.conv3d(xt, wt, bt, padding=1)
    out.backward(torch.from_numpy(np.ascontiguousarray(np.moveaxis(dout, -1, 0)))[None])
    dx = co.conv3d_backward_data(dout, wf)
    np.testing.assert_allclose(dx, np.moveaxis(xt.grad[0].numpy(), 0, -1), rtol=1e-10, atol=1e-9)
    dw, db = co.conv3d_weight_grad(x, dout)
    expected = wt.grad.numpy().transpose(2, 3, 4, 0, 1).reshape(co.TAPS, co.C, co.C)
    np.testing.assert_allclose(dw, expected, rtol=1e-10, atol=1e-8)
    np.testing.assert_allclose(db, bt.grad.numpy(), rtol=1e-12)


def test_network_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    shape = (3, 3, 4)
    layers = 2
    params = rng.normal(scale=0.05, size=layers * co.LAYER_PARAMS)
    features = rng.normal(size=shape + (co.C,))
    target = rng.normal(size=shape)
    weight = rng.random(shape)
    loss, grad, _ = co.network_gradient(features, params, target, weight, emulate_bf16=False)
    picks = rng.choice(len(params), 12, replace=False)
    picks = np.concatenate([picks, [co.TAPS * co.C * co.C + 3, co.LAYER_PARAMS + 13 * co.C * co.C + 5]])
    for p in picks:
        h = 1e-6
        up, down = params.copy(), params.copy()
        up[p] += h
        down[p] -= h
        # FP32 inputs are not rounded without emulation, so the loss is smooth in the parameter
        lu = co.network_gradient(features, up, target, weight, emulate_bf16=False)[0]
        ld = co.network_gradient(features, down, target, weight, emulate_bf16=False)[0]
        numeric = (lu - ld) / (2 * h)
        assert grad[p] == pytest.approx(numeric, rel=2e-3, abs=1e-6), p


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.normal(size=4096).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 4096),
                        np.array([0.0, -0.0, 1.0, 1.00390625, 1.01171875, 3.0e38, -1e-40], np.float32)])
    expected = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(co.bf16_round(x), expected)
    assert np.array_equal(co.from_bf16_bits(co.bf16_bits(x)), expected)


def test_adam_matches_torch():
    rng = np.random.default_rng(9)
    w = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(3)]
    p = torch.tensor(w, requires_grad=True)
    opt = torch.optim.Adam([p], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    m, v, master = np.zeros(50), np.zeros(50), w.copy()
    for step, g in enumerate(grads, start=1):
        p.grad = torch.tensor(g)
        opt.step()
        master, m, v = co.adam(master, m, v, g, 1e-3, 0.9, 0.999, 1e-8, step)
    np.testing.assert_allclose(master, p.detach().numpy(), rtol=1e-12, atol=1e-14)
