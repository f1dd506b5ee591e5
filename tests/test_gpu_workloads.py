"""Full-size workloads on the GPU path.

* 2-D Euler isentropic vortex, 20x20 quads on [-10,10]^2, P=3, DG, Rusanov, RK45,
  dt = 0.005, 200 steps: parity with the oracle run over the same 200 steps, and
  the error against the analytically advected vortex.
* 3-D Navier-Stokes Taylor-Green, 32^3 hexes on [0,2pi]^3, P=3, Re 1600, Mach 0.1:
  size-independent properties (free-stream preservation, discrete conservation)
  and parity with the oracle for one residual and one RK45 step.
"""
import math
import os

import numpy as np
import pytest

import cases
import frmesh
import hifr
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TWO_PI = cases.TWO_PI


def _vortex_case():
    L = 20.0
    xv, conn = frmesh.box_quad(20, 20, L, L, origin=(-10.0, -10.0))
    c2v = frmesh.gambit_to_internal(1, conn)
    deck = hifr.deck_text(viscous=0, order=3, vcjh_scheme_quad=1, dt=0.005, adv_type=3, dx_cyclic=L, dy_cyclic=L)
    s = hifr.Solver.from_arrays(deck, 1, xv, c2v)
    od = po.default_deck(equation=0, viscous=0, order=3, adv_type=3, dt=0.005, vcjh_scheme=1, dx_cyclic=L, dy_cyclic=L)
    o = po.Oracle(od, 1, xv, c2v, threads=min(16, os.cpu_count() or 1))
    return s, o


def test_isentropic_vortex_200_steps(require_gpu):
    s, o = _vortex_case()
    pos = o.get("pos_upts")
    u0 = frmesh.shu_vortex_state(pos)
    s.set_state(u0)
    o.set_state(u0)
    o.zero_rk_register()
    s.run(200)
    o.run_steps(200)
    us, uo = s.get("disu_upts"), o.get("disu_upts")
    assert cases.field_rel_err(us, uo).max() < 1e-10
    # against the exact solution: vortex advected by u_inf * t (t = 1)
    exact = frmesh.shu_vortex_state(pos, t=1.0)
    err = np.abs(us - exact).max(axis=(0, 1))
    # 20x20 P=3 on [-10,10]^2 resolves the unit-radius core with ~1 element per radius
    assert err.max() < 1.5e-2, err


def _tgv_solver(n=32, order=3, host_only=False, metrics=0):
    deck = hifr.deck_text(viscous=1, order=order, vcjh_scheme_hexa=1, dt=1e-3, adv_type=3, tau=1.0, pen_fact=0.5,
                          gamma=1.4, prandtl=0.71, Mach_free_stream=0.1, nx_free_stream=1.0,
                          Re_free_stream=1600.0, L_free_stream=1.0, T_free_stream=300.0, fix_vis=1,
                          dx_cyclic=TWO_PI, dy_cyclic=TWO_PI, dz_cyclic=TWO_PI, hifr_host_only=int(host_only),
                          hifr_metrics=metrics)
    return hifr.Solver.box(deck, hifr.HEX, (n, n, n), (TWO_PI,) * 3, (0.0,) * 3), deck


def _tgv_state(pos, seed=1234):
    u = frmesh.tgv_state(pos, p0=1.0 / (1.4 * 0.1 ** 2))
    return u * (1.0 + frmesh.seeded_perturbation(u.shape, 1e-3, seed))


def test_tgv_32cubed_free_stream_and_conservation(require_gpu):
    s, _ = _tgv_solver()
    pos = s.get("pos_upts")
    u = np.zeros((s.nu, s.n_eles, s.nf))
    u[..., 0] = 1.0
    u[..., 1] = 0.3
    u[..., 2] = -0.2
    u[..., 3] = 0.1
    u[..., 4] = 71.4 / 0.4 + 0.5 * (0.09 + 0.04 + 0.01)
    s.set_state(u)
    div = s.residual()
    assert np.abs(div).max() < 1e-10
    s.set_state(_tgv_state(pos))
    div = s.residual()
    w1 = np.polynomial.legendre.leggauss(4)[1]
    wu = np.einsum("i,j,k->kji", w1, w1, w1).ravel()
    total = np.einsum("j,jek->k", wu, div)
    scale = np.einsum("j,jek->k", wu, np.abs(div))
    assert np.all(np.abs(total) < 1e-12 * scale + 1e-12), (total, scale)


# Residual tolerance per conserved variable, relative to max|R| of that variable.
#  * per_point: the context keeps the reference's per-point metrics; the oracle
#    uses the same arrays: 1e-12.
#  * per_element: the context keeps one JGinv / detJ per element and one normal /
#    tdA per face (the box is affine); the oracle is given exactly those metrics
#    (downloaded from the context, expanded to the per-point layouts): 1e-12.
#  * per_element_vs_reference_metrics: the same context against the oracle's own
#    per-point metrics, which are different inputs: the reference's per-point
#    JGinv values differ between the points of one element by rounding (~1e-16
#    relative); at Mach 0.1 the momentum flux carries the mean pressure
#    p0 = 71.4, so that rounding enters the momentum divergence as
#    ~p0 * |D| * 1e-16 ~ 5e-14 absolute, ~1e-11 of the momentum residual.
RESID_TOL = {"per_point": [1e-12] * 5, "per_element": [1e-12] * 5,
             "per_element_vs_reference_metrics": [1e-12, 3e-11, 3e-11, 3e-11, 1e-12]}


@pytest.mark.parametrize("case", ["per_point", "per_element", "per_element_vs_reference_metrics"])
def test_tgv_32cubed_against_oracle(require_gpu, case):
    metrics = 1 if case == "per_point" else 0
    s, _ = _tgv_solver(metrics=metrics)
    assert s.metrics_per_element() == (metrics == 0)
    xv = s.get("xv").reshape(-1, 3)
    c2v = s.get("c2v").reshape(-1, 8).astype(np.int32)
    od = po.default_deck(equation=0, viscous=1, order=3, adv_type=3, dt=1e-3, tau=1.0, pen_fact=0.5,
                         gamma=1.4, prandtl=0.71, Mach_free_stream=0.1, nx_free_stream=1.0,
                         Re_free_stream=1600.0, L_free_stream=1.0, T_free_stream=300.0, vcjh_scheme=1,
                         dx_cyclic=TWO_PI, dy_cyclic=TWO_PI, dz_cyclic=TWO_PI)
    o = po.Oracle(od, 4, xv, c2v, threads=min(16, os.cpu_count() or 1))
    if case == "per_element":
        dm = s.device_metrics()
        # the context's per-element values are the reference's at each element's
        # first point, to rounding of the metric's scale (the box mesh's zero
        # entries are ~1e-17 at some points of the reference's per-point values)
        jd, jo = dm["JGinv_upts"], o.get("JGinv_upts")
        assert np.abs(jd - jo).max() < 1e-13 * np.abs(jo).max()
        o.set_metrics(dm["detjac_upts"], dm["JGinv_upts"], dm["tdA_fpts"], dm["norm_fpts"])
    u = _tgv_state(o.get("pos_upts"))
    s.set_state(u)
    o.set_state(u)
    o.residual()
    err = cases.field_rel_err(s.residual(), o.get("div_tconf_upts"))
    print(case, "residual field errors", err)
    assert np.all(err < RESID_TOL[case]), err
    s.set_state(u)
    o.set_state(u)
    o.zero_rk_register()
    s.run(1)
    o.run_steps(1)
    # the z-momentum starts at zero (w = 0), so after one step its scale is
    # dt * |R|: the state carries the residual's tolerance
    err = cases.field_rel_err(s.get("disu_upts"), o.get("disu_upts"))
    print(case, "state field errors after one RK45 step", err)
    assert np.all(err < RESID_TOL[case]), err
