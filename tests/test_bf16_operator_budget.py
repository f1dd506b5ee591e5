"""Error budget of BF16 extrapolation operators on the mixed path.

Mixed precision keeps the per-variable error within 1e-6 of FP64.  This file
measures what rounding the solution-to-flux-point extrapolation operators to
BF16 would cost, on the 3-D Taylor-Green field (Mach 0.1) sampled at the Gauss
points of 8^3 hexes at P=3 and extrapolated along lines to both faces, with
FP32 arithmetic on deviations from the mean state as the mixed kernels do it:

* FP32 operators stay within the budget (5e-8 on momentum);
* BF16 operators miss it by three orders of magnitude (2.5e-3), and even a
  two-term BF16 split (hi + lo, two MFMA products) misses it (8.6e-6);
* BF16 weights no longer sum to one, so a uniform state is not preserved.

The hex extrapolation is a length-(P+1) sum per flux point on the VALU, where
BF16 operands would save no instruction, so the mixed kernels keep FP32
operators.  The 1-D Lagrange factors at +-1 over the Gauss points are the
factors of the reference's hex nodal basis (src/eles_hexas.cpp:917-930 with
eval_lagrange, src/funcs.cpp:57-74), evaluated on the faces.
"""
import numpy as np
import pytest
import torch

import frmesh

P = 3
BUDGET = 1e-6


def _lagrange_at(xs, x):
    return np.array([np.prod([(x - xs[j]) / (xs[i] - xs[j]) for j in range(len(xs)) if j != i])
                     for i in range(len(xs))])


def _bf16(w):
    return torch.tensor(np.asarray(w, np.float32)).to(torch.bfloat16).to(torch.float64).numpy()


def _f32(w):
    return np.asarray(w, np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def lines():
    """TGV state on lines of Gauss points along x: (elements, points, y, z, 5)."""
    n = P + 1
    xg, _ = np.polynomial.legendre.leggauss(n)
    ne, h = 8, 2.0 * np.pi / 8
    xs = (np.arange(ne)[:, None] * h + (xg[None, :] + 1.0) * 0.5 * h).reshape(-1)
    X, Y, Z = np.meshgrid(xs, xs, xs, indexing="ij")
    u = frmesh.tgv_state(np.stack([X, Y, Z], -1), p0=1.0 / (1.4 * 0.1 ** 2))
    return xg, u.reshape(ne, n, ne * n, ne * n, 5)


def _trace_error(xg, u, weights):
    """Relative L2 error per variable of FP32 extrapolation with the given weights
    (one array per face) against FP64 with the exact weights."""
    ref = np.array([u[..., 0].mean(), 0.0, 0.0, 0.0, u[..., 4].mean()])
    dev = u - ref
    errs = []
    for exact_w, w in zip((_lagrange_at(xg, -1.0), _lagrange_at(xg, 1.0)), weights):
        exact = np.einsum("q,eqyzk->eyzk", exact_w, u)
        acc = np.zeros_like(exact)
        for q in range(len(xg)):
            acc = _f32(acc + _f32(_f32(w[q]) * _f32(dev[:, q])))
        errs.append(((acc + ref - exact) ** 2).reshape(-1, 5).sum(0) / (exact ** 2).reshape(-1, 5).sum(0))
    return np.sqrt(np.sum(errs, axis=0) / 2.0)


def _faces(xg):
    return _lagrange_at(xg, -1.0), _lagrange_at(xg, 1.0)


def test_fp32_operators_meet_the_budget(lines):
    xg, u = lines
    e = _trace_error(xg, u, [_f32(w) for w in _faces(xg)])
    assert np.all(e[[0, 1, 2, 4]] < BUDGET / 10), e


def test_bf16_operators_miss_the_budget(lines):
    xg, u = lines
    e = _trace_error(xg, u, [_bf16(w) for w in _faces(xg)])
    assert e[1] > 1000 * BUDGET and e[2] > 1000 * BUDGET, e


def test_split_bf16_operators_miss_the_budget(lines):
    xg, u = lines
    e = _trace_error(xg, u, [_bf16(w) + _bf16(w - _bf16(w)) for w in _faces(xg)])
    assert e[1] > 5 * BUDGET and e[2] > 5 * BUDGET, e


def test_bf16_weights_break_the_partition_of_unity():
    xg, _ = np.polynomial.legendre.leggauss(P + 1)
    for w in _faces(xg):
        assert abs(w.sum() - 1.0) < 1e-14
        assert abs(_bf16(w).sum() - 1.0) > 1e-3
