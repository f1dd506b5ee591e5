"""Roe interface flux (riemann_solve_type 2) on the device in 3-D: hexes and tets,
Euler and Navier-Stokes, periodic faces and isothermal walls, per-point and
per-element metrics, against the oracle on the same mesh and state.

The 2-D Roe flux is compared with the reference's own residuals and RK45
trajectories by the golden tests of test_gpu_parity.py (quad_*_roe_* fixtures);
the reference stops in 3-D, so here the oracle's 3-D form, pinned by the
reduction and rotation checks of test_oracle_properties.py, is the comparison.
Tolerance: 1e-12 of max|R| per conserved variable; RK45 states 1e-12.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

TOL = 1e-12

# (fixture whose mesh and state are reused, deck overrides)
CASES = [
    ("hex_p3_ns_perturbed_3x3x3", {}),
    ("hex_p2_euler_3x2x2", {}),
    ("hex_p3_channel_walls_3x3x2", {}),
    ("hex_p3_ns_sutherland_perturbed_2x2x2", {}),
    ("tet_p3_ns_perturbed_2x2x2", {}),
    ("tet_p2_euler_2x2x2", {}),
    ("tet_p3_channel_wale_walls_2x2x2", {"LES": 0}),
]


def _solver_and_oracle(g, metrics, over):
    s = cases.solver_of(g, riemann_solve_type=2, hifr_metrics=metrics, **over)
    o = cases.oracle_of(g, riemann_solve_type=2, **over)
    return s, o


@pytest.mark.parametrize("metrics", [0, 1], ids=["auto", "per_point"])
@pytest.mark.parametrize("name,over", CASES, ids=[c[0] for c in CASES])
def test_roe_residual_matches_oracle(require_gpu, name, over, metrics):
    g = cases.load_golden(name)
    print("case", name, "metrics", metrics, flush=True)
    s, o = _solver_and_oracle(g, metrics, over)
    s.set_state(g["u0"])
    div = s.residual()
    o.set_state(g["u0"])
    o.residual()
    assert cases.field_rel_err(div, o.get("div_tconf_upts")).max() < TOL
    # the Roe residual differs from the Rusanov one (the flux switch took effect)
    r = cases.solver_of(g, hifr_metrics=metrics, **over)
    r.set_state(g["u0"])
    assert cases.field_rel_err(r.residual(), div).max() > 1e-8


@pytest.mark.parametrize("name,over", [CASES[0], CASES[2], CASES[6]], ids=[CASES[i][0] for i in (0, 2, 6)])
def test_roe_rk45_matches_oracle(require_gpu, name, over):
    g = cases.load_golden(name)
    print("case", name, flush=True)
    s, o = _solver_and_oracle(g, 0, dict(over, adv_type=3, dt=1e-3))
    s.set_state(g["u0"])
    s.run(3)
    o.set_state(g["u0"])
    o.zero_rk_register()
    o.run_steps(3)
    assert cases.field_rel_err(s.get("disu_upts"), o.get("disu_upts")).max() < TOL


def test_roe_interface_flux_is_conservative(require_gpu):
    """Both sides of an interface evaluate the Roe flux in one orientation, so on a
    periodic mesh the total of the residual integrated over the domain vanishes to
    rounding: sum over elements and points of w_j div_j = 0 (the divergence
    weighted by the Gauss weights is the surface integral of the common flux)."""
    g = cases.load_golden("hex_p3_ns_perturbed_3x3x3")
    s = cases.solver_of(g, riemann_solve_type=2)
    s.set_state(g["u0"])
    div = s.residual()
    w1 = np.polynomial.legendre.leggauss(int(g["order"]) + 1)[1]
    w3 = np.einsum("i,j,k->ijk", w1, w1, w1).ravel()
    tot = np.einsum("j,jek->k", w3, div)
    scale = np.einsum("j,jek->k", w3, np.abs(div))
    assert np.all(np.abs(tot) < 1e-12 * scale), tot / scale


def test_roe_rejected_where_not_built(require_gpu):
    g = cases.load_golden("hex_p3_les_wale_perturbed_3x3x3")
    with pytest.raises(RuntimeError, match="unsupported"):
        cases.solver_of(g, riemann_solve_type=2)
    g = cases.load_golden("hex_p3_ns_perturbed_3x3x3")
    with pytest.raises(RuntimeError, match="unsupported"):
        cases.solver_of(g, riemann_solve_type=2, hifr_precision=1)
    with pytest.raises(RuntimeError, match="unsupported"):
        cases.solver_of(g, riemann_solve_type=1)
