"""The oracle against the reference's own numbers.

* tests/golden/*.npz were written by tests/golden/make_goldens.py from the
  reference solver compiled from its sources (oracle/Makefile.ref): the residual
  div/J recovered from one forward-Euler step, and RK45 states, on seeded states
  and periodic meshes (affine and perturbed hexes, quads, P = 2..4, Euler and
  Navier-Stokes, DG / SD / Hu correction functions).  The reference prints 15
  significant digits, so agreement is checked to that precision.
* tgv_regression.json holds the reference's regression vector
  (testcases/regression_tests.py:300-307): L1 residual norms after 25 RK45 steps
  of the Taylor-Green deck on the 15^3 mesh, absolute tolerance 1e-5.
"""
import json
import os

import numpy as np
import pytest

import cases
import frmesh
from oracle import pyoracle as po


def _oracle(g, **over):
    return cases.oracle_of(g, **over)


@pytest.mark.parametrize("name", cases.golden_names("residual"))
def test_oracle_residual_matches_reference(name):
    g = cases.load_golden(name)
    o = _oracle(g)
    o.set_state(g["u0"])
    o.residual()
    R = o.get("div_tconf_upts") / o.get("detjac_upts")[:, :, None]
    rel = cases.field_rel_err(R, g["R"])
    if int(g["ele_type"]) == 2:
        # tets: the reference interpolates the injected restart state through its
        # nodal basis at the same points, V V^-1 = I + O(1e-15), which moves its
        # residual by ~1e-13 of max|R|; the operators themselves agree to rounding
        assert rel.max() < 2e-13, rel
        return
    # R is recovered from 15-digit restart output: ~1e-15 of max|R|
    assert rel.max() < 2e-14, rel
    # the forward-Euler state of the oracle printed with 15 digits equals the
    # reference's printed state digit for digit
    o.advance(0, 0)
    u1 = o.get("disu_upts")
    u1_ref = g["u0"] - float(g["dt"]) * g["R"]
    same = np.mean(np.isclose(u1, u1_ref, rtol=5e-14, atol=0))
    assert same == 1.0


@pytest.mark.parametrize("name", cases.golden_names("rk45"))
def test_oracle_rk45_matches_reference(name):
    g = cases.load_golden(name)
    o = _oracle(g)
    o.set_state(g["u0"])
    o.zero_rk_register()
    o.run_steps(int(g["steps"]))
    u = o.get("disu_upts")
    if int(g["ele_type"]) == 2:
        # tets: see the residual test (restart interpolation at 1e-15)
        assert cases.field_rel_err(u, g["u_final"]).max() < 1e-13
        return
    printed = np.array([float("%.15g" % v) for v in u.ravel(order="F")]).reshape(u.shape, order="F")
    assert np.array_equal(printed, g["u_final"])


def _sutherland_goldens():
    return [n for n in cases.golden_names() if int(cases.load_golden(n).get("fix_vis", 1)) == 0]


def test_sutherland_goldens_exist():
    assert len(_sutherland_goldens()) >= 4


@pytest.mark.parametrize("name", _sutherland_goldens())
def test_sutherland_goldens_pin_the_law(name):
    """The fix_vis 0 fixtures depend on Sutherland's law: the oracle with the law
    switched off (fix_vis 1, constant viscosity mu_inf) misses them by far more
    than their 15-digit precision, so a wrong exponent or constant in the
    oracle's law (or in the non-dimensionalisation that sets c_sth, rt_inf and
    mu_inf) fails the reference comparisons above."""
    g = cases.load_golden(name)
    o = _oracle(g, fix_vis=1.0)
    o.set_state(g["u0"])
    if str(g["kind"]) == "residual":
        o.residual()
        R = o.get("div_tconf_upts") / o.get("detjac_upts")[:, :, None]
        rel = cases.field_rel_err(R, g["R"])
        assert rel[1:].max() > 1e-6, rel
    else:
        o.zero_rk_register()
        o.run_steps(int(g["steps"]))
        rel = cases.field_rel_err(o.get("disu_upts"), g["u_final"])
        assert rel[1:].max() > 1e-9, rel


@pytest.mark.slow
def test_oracle_tgv_regression_norms():
    with open(os.path.join(cases.GOLDEN, "tgv_regression.json")) as fh:
        reg = json.load(fh)
    m, d = reg["mesh"], reg["deck"]
    L = m["length"]
    xv, conn = frmesh.box_hex(*m["n"], L, L, L, digits=m["digits"])
    c2v = frmesh.gambit_to_internal(4, conn)
    deck = po.default_deck(equation=0, viscous=1, order=d["order"], adv_type=3, dt=d["dt"], tau=d["tau"],
                           pen_fact=d["pen_fact"], ic_form=d["ic_form"], gamma=d["gamma"],
                           prandtl=d["prandtl"], Mach_free_stream=d["Mach_free_stream"],
                           nx_free_stream=d["nx_free_stream"], Re_free_stream=d["Re_free_stream"],
                           L_free_stream=d["L_free_stream"], T_free_stream=d["T_free_stream"],
                           vcjh_scheme=d["vcjh_scheme_hexa"], dx_cyclic=L, dy_cyclic=L, dz_cyclic=L)
    o = po.Oracle(deck, 4, xv, c2v, threads=os.cpu_count())
    o.set_ics_tgv(1.0)
    o.run_steps(1)
    np.testing.assert_allclose(o.norms(1), reg["step1_vals"], atol=reg["tol"], rtol=0)
    o.run_steps(reg["test_iter"] - 1)
    np.testing.assert_allclose(o.norms(1), reg["test_vals"], atol=reg["tol"], rtol=0)
