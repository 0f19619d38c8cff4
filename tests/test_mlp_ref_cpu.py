"""Pins the FP32 reference of the surrogate kernels (tests/mlp_ref.py) on the
CPU: the manual backward against torch autograd, GELU's derivative against a
central difference, and the padding invariants the kernels rely on."""
import math

import torch

import mlp_ref as R


def test_gelu_matches_torch():
    z = torch.linspace(-6, 6, 1001, dtype=torch.float64)
    assert torch.allclose(R.gelu(z), torch.nn.functional.gelu(z), atol=1e-12)


def test_gelu_grad_central_difference():
    z = torch.linspace(-5, 5, 401, dtype=torch.float64)
    h = 1e-5
    fd = (R.gelu(z + h) - R.gelu(z - h)) / (2 * h)
    assert torch.allclose(R.gelu_grad(z), fd, atol=1e-9)


def test_manual_backward_matches_autograd():
    Ws, bs = R.init_weights(3)
    x, t = R.random_inputs(4, 64)
    Ws = [w.double().requires_grad_() for w in Ws]
    bs = [b.double().requires_grad_() for b in bs]
    x, t = x.double(), t.double()
    y, cache = R.forward(Ws, bs, x, rounded=False)
    loss, dy = R.loss_and_grad_out(y, t)
    loss.backward()
    with torch.no_grad():
        dWs, dbs = R.backward(Ws, cache, dy, rounded=False)
    for l in range(5):
        assert torch.allclose(dWs[l], Ws[l].grad, rtol=1e-10, atol=1e-14), l
        assert torch.allclose(dbs[l], bs[l].grad, rtol=1e-10, atol=1e-14), l


def test_padding_gradients_are_zero():
    Ws, bs = R.init_weights(5)
    x, t = R.random_inputs(6, 32)
    y, cache = R.forward(Ws, bs, x, rounded=True)
    assert torch.all(y[:, 6:] == 0)
    _, dy = R.loss_and_grad_out(y[:, :], t)
    dWs, dbs = R.backward(Ws, cache, dy, rounded=True)
    assert torch.all(dWs[0][:, 13:] == 0)
    assert torch.all(dWs[4][6:] == 0) and torch.all(dbs[4][6:] == 0)


def test_rounded_forward_close_to_exact():
    # BF16 rounding moves the output by well under a percent of its scale
    Ws, bs = R.init_weights(7)
    x, _ = R.random_inputs(8, 256)
    y1, _ = R.forward(Ws, bs, x, rounded=True)
    y0, _ = R.forward(Ws, bs, x, rounded=False)
    rel = (y1 - y0).norm() / y0.norm()
    assert rel < 2e-2 and math.isfinite(float(rel))
