"""The learned-SGS network's device kernels (BF16 MFMA forward/backward, Adam)
against the oracle (oracle/sgs_mlp.py, which replays the same BF16 roundings with
FP64 accumulation).

Tolerances: the device accumulates in FP32 and its exp differs from the oracle's
in the last bits, so an activation or output gradient occasionally rounds to the
neighbouring BF16 value (2^-8 relative); gradients agree to 2e-3 norm-wise per
parameter tensor, outputs to 1e-2 of max|y| and 1e-3 on average.
"""
import numpy as np
import pytest

import hifr
from oracle import sgs_mlp as om

pytestmark = pytest.mark.gpu


def _data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, om.N_IN)).astype(np.float32)
    t = rng.uniform(0.0, 2.0, n).astype(np.float32)
    return x, t


def _split(g):
    out, o = {}, 0
    for name, shp in om.SHAPES:
        k = int(np.prod(shp))
        out[name] = np.asarray(g[o:o + k], dtype=np.float64)
        o += k
    return out


@pytest.mark.parametrize("n", [16, 1000, 70001])
def test_loss_gradient_matches_oracle(require_gpu, n):
    p = om.init_params(21)
    x, t = _data(n, 22)
    m = hifr.MLP()
    m.set_params(p)
    loss, g = m.loss_grad(x, t)
    lo, go = om.loss_grad(p, x, t, rounding=True)
    print("n", n, "loss", loss, lo, flush=True)
    assert abs(loss - lo) <= 1e-5 * lo
    gs, gos = _split(g), _split(go)
    for name, _ in om.SHAPES:
        err = np.linalg.norm(gs[name] - gos[name]) / np.linalg.norm(gos[name])
        print("  %s rel err %.2e" % (name, err), flush=True)
        assert err < 2e-3, (name, err)


def test_predict_matches_oracle(require_gpu):
    p = om.init_params(31)
    x, _ = _data(4099, 32)
    m = hifr.MLP()
    m.set_params(p)
    y = m.predict(x)
    yo = om.forward(p, x, rounding=True)
    d = np.abs(y - yo)
    print("predict max %.2e mean %.2e (max|y| %.2f)" % (d.max(), d.mean(), np.abs(yo).max()), flush=True)
    assert d.max() < 1e-2 * np.abs(yo).max() and d.mean() < 1e-3 * np.abs(yo).mean()


def test_params_round_trip(require_gpu):
    p = om.init_params(41)
    m = hifr.MLP()
    m.set_params(p)
    assert np.array_equal(m.params(), p)


def test_one_adam_step_matches_oracle(require_gpu):
    """One full-batch step: the device update equals the oracle's Adam step with
    the oracle's gradient, except where the gradient is so small that the two
    gradients differ in sign (Adam's first step is +-lr for any nonzero g)."""
    p = om.init_params(51)
    x, t = _data(512, 52)
    m = hifr.MLP()
    m.set_params(p)
    m.set_data(x, t)
    m.train(1, 512, 1e-3)
    p1 = m.params().astype(np.float64)
    _, go = om.loss_grad(p, x, t, rounding=True)
    po, _, _ = om.adam_step(p.astype(np.float64), 0.0, 0.0, go, 1, 1e-3)
    bad = np.abs(p1 - po) > 1e-6
    print("params differing by > 1e-6: %d of %d" % (bad.sum(), bad.size), flush=True)
    assert bad.mean() < 5e-3


def test_training_reduces_loss(require_gpu):
    """Adam on a smooth synthetic regression: the device-trained network fits it."""
    rng = np.random.default_rng(61)
    n = 65536
    x = rng.standard_normal((n, om.N_IN)).astype(np.float32)
    t = (0.5 * (1.0 + np.tanh(x[:, 0] - 0.5 * x[:, 1])) + 0.1 * x[:, 2] ** 2).astype(np.float32)
    m = hifr.MLP(seed=7)
    m.set_data(x, t)
    hist = m.train(20, 1024, 2e-3)
    print("epoch losses", hist.tolist(), flush=True)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.1 * hist[0]
    # held-out samples
    xv = rng.standard_normal((4096, om.N_IN)).astype(np.float32)
    tv = 0.5 * (1.0 + np.tanh(xv[:, 0] - 0.5 * xv[:, 1])) + 0.1 * xv[:, 2] ** 2
    yv = m.predict(xv)
    rel = np.sqrt(np.mean((yv - tv) ** 2) / np.mean((tv - tv.mean()) ** 2))
    print("held-out relative RMS error %.3f" % rel, flush=True)
    assert rel < 0.2
