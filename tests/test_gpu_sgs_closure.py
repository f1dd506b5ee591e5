"""The learned LES closure on the GPU against the oracle: the training samples
(modal filter, SGS stress, exact eddy viscosity) and the residual of the
HIFR_SGS_LEARNED kernels with and without a network attached.

Tolerances: samples are FP32 outputs of FP64 arithmetic (1e-5 of the largest
value); the residual with the WALE fallback everywhere is FP64 (1e-12 of max|R|
per variable, as every FP64 residual); with a network, the eddy viscosity passes
through BF16 operands, and an input that rounds to the neighbouring BF16 value on
one side changes nu_t there by up to 2^-8 relative: 1e-4.
"""
import numpy as np
import pytest

import cases
import frmesh
import hifr
from oracle import sgs_closure as sc
from oracle import sgs_mlp as om

pytestmark = pytest.mark.gpu

TWO_PI = cases.TWO_PI


def _case(order, n=3, filter_ratio=2.0, noise=2e-2, seed=5):
    xv, conn = frmesh.box_hex(n, n, n, TWO_PI, TWO_PI, TWO_PI)
    g = dict(ele_type=4, xv=xv, conn_gambit=conn, order=order, vcjh=1, viscous=1, tau=1.0, pen_fact=0.5,
             cyc=np.array([TWO_PI] * 3), dt=1e-3, kind="residual", les=1, sgs_model=1, filter_ratio=filter_ratio,
             Re=5000.0)
    return g


def _state(pos, noise=2e-2, seed=5):
    u = frmesh.tgv_state(pos, p0=1.0 / (1.4 * 0.1 ** 2))
    return u * (1.0 + frmesh.seeded_perturbation(u.shape, noise, seed))


def test_samples_match_oracle(require_gpu):
    g = _case(5, n=3)
    s = cases.solver_of(g, LES=0)
    u0 = _state(s.get("pos_upts"))
    s.set_state(u0)
    x, ce, cw, ulo = s.sgs_samples(2, filter_ratio=1.0, want_state=True)
    mu = cases.oracle_of(g, LES=0).phys().mu_inf
    # dr_d/dx_m = JGinv(d, m) / detJ of each (affine) element, from its first point
    JG = s.get("JGinv_upts")[:, :, 0, :]        # (d, m, N)
    detj = s.get("detjac_upts")[0, :]
    drdx = JG.transpose(2, 0, 1) / detj[:, None, None]
    xo, ceo, cwo, ulo_o = sc.samples(u0, hifr.gauss_points(5), hifr.gauss_points(2), drdx, detj, 1.0, mu)
    print("state", np.abs(ulo - ulo_o).max(), "x", np.abs(x - xo).max(), "c_exact", np.abs(ce - ceo).max(),
          np.abs(ceo).max(), "c_wale", np.abs(cw - cwo).max(), np.abs(cwo).max(), flush=True)
    assert np.abs(ulo - ulo_o).max() < 1e-10 * np.abs(ulo_o).max()
    assert np.abs(x - xo).max() < 1e-5 * np.abs(xo).max()
    assert np.abs(ce - ceo).max() < 1e-5 * np.abs(ceo).max()
    assert np.abs(cw - cwo).max() < 1e-5 * np.abs(cwo).max()


def _oracle_residual(g, u0, params=None, norm=None):
    o = cases.oracle_of(g, SGS_model=10)
    o.set_state(u0)
    o.residual()
    grad = o.get("grad_disu_upts")          # (Nu, N, nf, d)
    nu_pts, N = grad.shape[0], grad.shape[1]
    u = np.asarray(u0).transpose(1, 0, 2).reshape(-1, 5)
    gg = grad.transpose(1, 0, 2, 3).reshape(-1, 5, 3)
    detj = o.get("detjac_upts").T.reshape(-1)
    delta = float(g["filter_ratio"]) * (detj * 8.0) ** (1.0 / 3.0) / (int(g["order"]) + 1)
    mu = o.phys().mu_inf
    nu = sc.closure_nu(u, gg, delta, mu, params, norm)
    o.set_sgs_nu(nu.reshape(N, nu_pts).T)
    o.residual()
    return o.get("div_tconf_upts"), nu


@pytest.mark.parametrize("order", [2, 3])
def test_wale_fallback_residual_matches_oracle(require_gpu, order):
    g = _case(order)
    s = cases.solver_of(g, hifr_learned_sgs=1)
    u0 = _state(s.get("pos_upts"))
    s.set_state(u0)
    R = s.residual()
    Ro, nu = _oracle_residual(g, u0)
    err = cases.field_rel_err(R, Ro)
    print("P", order, "WALE fallback: field errors", err, "max nu_t", nu.max(), flush=True)
    assert nu.max() > 0
    assert np.all(err < 1e-12), err


def _network(seed=3):
    m = hifr.MLP()
    p = om.init_params(seed)
    p[om.N_PARAMS - 1] = -3.0  # softplus(-3) ~ 0.05: a modest coefficient
    m.set_params(p)
    mean = np.zeros(10, np.float32)
    mean[9] = 5.0
    scale = np.ones(10, np.float32)
    scale[9] = 2.0
    lo = np.full(10, -1e30, np.float32)
    hi = np.full(10, 1e30, np.float32)
    m.set_normalization(mean, scale, lo, hi, 0.1)
    return m, p, (mean, scale, lo, hi, 0.1)


@pytest.mark.parametrize("order", [2, 3])
def test_network_residual_matches_oracle(require_gpu, order):
    g = _case(order)
    s = cases.solver_of(g, hifr_learned_sgs=1)
    m, p, norm = _network()
    s.set_sgs_network(m)
    u0 = _state(s.get("pos_upts"))
    s.set_state(u0)
    R = s.residual()
    Ro, nu = _oracle_residual(g, u0, p, norm)
    Rw, _ = _oracle_residual(g, u0)
    err = cases.field_rel_err(R, Ro)
    dw = cases.field_rel_err(Rw, Ro)
    print("P", order, "network: field errors", err, "(network vs WALE residual differ by", dw, ")", flush=True)
    assert np.all(dw[1:] > 1e-6)  # the network's eddy viscosity is not WALE's (no SGS density flux)
    assert np.all(err < 1e-4), err


def test_network_weights_are_copied_on_attach(require_gpu):
    """The context keeps its own copy of the attached weights: closing the network
    (freeing its device buffers) or changing its parameters afterwards leaves the
    residual unchanged; detaching gives the WALE fallback again."""
    g = _case(2)
    s = cases.solver_of(g, hifr_learned_sgs=1)
    u0 = _state(s.get("pos_upts"))
    s.set_state(u0)
    R_wale = s.residual().copy()
    m, p, norm = _network()
    s.set_sgs_network(m)
    R_net = s.residual().copy()
    q = p.copy()
    q[om.N_PARAMS - 1] = 1.0
    m.set_params(q)
    R_changed = s.residual().copy()
    m.close()
    R_closed = s.residual().copy()
    s.set_sgs_network(None)
    R_detached = s.residual().copy()
    d = cases.field_rel_err(R_net, R_wale)
    print("network vs WALE", d, flush=True)
    assert np.all(d[1:] > 1e-6)
    assert np.array_equal(R_changed, R_net) and np.array_equal(R_closed, R_net)
    assert np.array_equal(R_detached, R_wale)


def test_mixed_learned_residual_close_to_oracle(require_gpu):
    g = _case(2)
    s = cases.solver_of(g, hifr_learned_sgs=1, hifr_precision=1)
    m, p, norm = _network()
    s.set_sgs_network(m)
    u0 = _state(s.get("pos_upts"))
    s.set_state(u0)
    R = s.residual()
    Ro, _ = _oracle_residual(g, u0, p, norm)
    err = cases.field_rel_err(R, Ro)
    print("mixed learned: field errors", err, flush=True)
    assert np.all(err < 1e-4), err
