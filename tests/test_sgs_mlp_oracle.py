"""Pins of the learned-SGS network oracle (oracle/sgs_mlp.py), CPU only.

- BF16 rounding: closed-form cases (ties to even) and torch's bfloat16 cast;
- silu, its derivative, softplus, sigmoid against torch in float64;
- with the BF16 roundings off, the loss gradient equals torch autograd in float64
  and central finite differences of the loss;
- one Adam step equals torch.optim.Adam.
A mistake in the backward pass (a missing silu' factor, a transposed weight, a
dropped 1/B) fails the autograd and finite-difference comparisons.
"""
import numpy as np
import pytest
import torch

from oracle import sgs_mlp as om


def test_bf16_closed_form():
    one = 1.0
    assert om.bf16(one) == 1.0
    assert om.bf16(1.0 + 2.0 ** -8) == 1.0                   # tie -> even (mantissa 0)
    assert om.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6   # tie -> even (mantissa 2)
    assert om.bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert om.bf16(-2.5) == -2.5


def test_bf16_matches_torch():
    x = np.random.default_rng(0).standard_normal(10000).astype(np.float32) * 10.0
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(om.bf16(x), ref)


def test_activations_match_torch():
    z = np.linspace(-30, 30, 1001)
    tz = torch.from_numpy(z)
    assert np.allclose(om.silu(z), torch.nn.functional.silu(tz).numpy(), rtol=1e-14, atol=1e-300)
    assert np.allclose(om.softplus(z), torch.nn.functional.softplus(tz).numpy(), rtol=1e-12, atol=1e-300)
    assert np.allclose(om.sigmoid(z), torch.sigmoid(tz).numpy(), rtol=1e-14)
    tz.requires_grad_(True)
    torch.nn.functional.silu(tz).sum().backward()
    assert np.allclose(om.d_silu(z), tz.grad.numpy(), rtol=1e-12, atol=1e-15)


def _torch_loss(p, x, t):
    w = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in om.unpack(p).items()}
    X = torch.tensor(np.asarray(x, dtype=np.float64))
    h = X.T
    for l in (1, 2, 3):
        h = torch.nn.functional.silu(w["W%d" % l] @ h + w["b%d" % l][:, None])
    y = torch.nn.functional.softplus(w["w4"] @ h + w["b4"][0])
    loss = ((y - torch.tensor(np.asarray(t, dtype=np.float64))) ** 2).mean()
    loss.backward()
    return float(loss.detach()), np.concatenate([w[name].grad.numpy().ravel() for name, _ in om.SHAPES])


def _data(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, om.N_IN)), rng.uniform(0.0, 2.0, n)


def test_gradient_matches_autograd():
    p = om.init_params(3)
    x, t = _data(37, 4)
    loss, g = om.loss_grad(p, x, t, rounding=False)
    tl, tg = _torch_loss(p, x, t)
    assert abs(loss - tl) < 1e-14 * max(1.0, tl)
    assert np.allclose(g, tg, rtol=1e-10, atol=1e-14)


def test_gradient_matches_finite_differences():
    p = om.init_params(5).astype(np.float64)
    x, t = _data(9, 6)
    _, g = om.loss_grad(p, x, t, rounding=False)
    rng = np.random.default_rng(7)
    idx = np.concatenate([rng.choice(om.N_PARAMS, 40, replace=False), [om.N_PARAMS - 1, 0, 640, 9024]])
    h = 1e-6
    for i in idx:
        e = np.zeros(om.N_PARAMS)
        e[i] = h
        lp = om.loss_grad(p + e, x, t, rounding=False)[0]
        lm = om.loss_grad(p - e, x, t, rounding=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_rounded_gradient_close_to_exact():
    """The BF16 roundings perturb the gradient by about the BF16 precision."""
    p = om.init_params(8)
    x, t = _data(64, 9)
    _, g0 = om.loss_grad(p, x, t, rounding=False)
    _, g1 = om.loss_grad(p, x, t, rounding=True)
    rel = np.linalg.norm(g1 - g0) / np.linalg.norm(g0)
    assert 1e-5 < rel < 5e-2


def test_adam_step_matches_torch():
    p = om.init_params(10).astype(np.float64)
    g = np.random.default_rng(11).standard_normal(om.N_PARAMS)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    tp = torch.tensor(p.copy(), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for step in (1, 2, 3):
        gs = g * step
        p, m, v = om.adam_step(p, m, v, gs, step, 1e-3)
        tp.grad = torch.tensor(gs)
        opt.step()
    assert np.allclose(p, tp.detach().numpy(), rtol=1e-12, atol=1e-15)


def test_pack_unpack_round_trip():
    p = om.init_params(12)
    assert p.size == om.N_PARAMS == 9089
    assert np.array_equal(om.pack(om.unpack(p)), p)
