"""Tets of every order (P = 1, 2, 3) on a small wall-bounded channel, against the
oracle: viscous, with and without LES (WALE and Smagorinsky), isothermal walls.

The tet kernels apply their operators as MFMA tiles whose shapes depend on the
order: 3 elements and one 16-column tile per wavefront at P = 3, 6 elements and
two tiles at P = 2 (zero-padded rows, Nu = 10), 16 elements and five tiles at
P = 1 (the 72-tet mesh leaves the last block half full).  Each case checks the
residual and two RK45 steps in FP64 (1e-12 of each field's maximum, like the
golden cases), the corrected-gradient traces (1e-12 of the largest component),
and the mixed path's residual (2e-5).
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("order,les,sgs", [(1, 0, 1), (1, 1, 1), (2, 0, 1), (2, 1, 1), (2, 1, 0), (3, 1, 0)],
                         ids=["p1", "p1_wale", "p2", "p2_wale", "p2_smag", "p3_smag"])
def test_tet_orders_against_oracle(require_gpu, order, les, sgs):
    g = cases.channel_tet_case(n=(3, 2, 2), order=order, perturb=0.1, dt=1e-4, steps=2, sgs_model=sgs)
    g["les"] = les
    print("tet case P=%d les=%d sgs_model=%d" % (order, les, sgs), flush=True)
    s, o = cases.solver_of(g), cases.oracle_of(g)
    u = cases.channel_tet_state(g, o.get("pos_upts"))
    s.set_state(u)
    o.set_state(u)
    o.residual()
    err = cases.field_rel_err(s.residual(), o.get("div_tconf_upts"))
    print("  residual", err, flush=True)
    assert err.max() < TOL
    # gradient traces, against the largest component: the channel's density is
    # uniform, so its gradient is zero up to roundoff and has no scale of its own
    gs, go = s.get("grad_disu_fpts"), o.get("grad_disu_fpts")
    gerr = np.abs(gs - go).max() / np.abs(go).max()
    print("  gradient traces %.2e" % gerr, flush=True)
    assert gerr < TOL
    o.set_state(u)
    o.zero_rk_register()
    o.run_steps(2)
    s.set_state(u)
    s.run(2)
    err = cases.field_rel_err(s.get("disu_upts"), o.get("disu_upts"))
    print("  2 RK45 steps", err, flush=True)
    assert err.max() < TOL
    # mixed path
    m = cases.solver_of(g, hifr_precision=1)
    m.set_state(u)
    o.set_state(u)
    o.residual()
    err = cases.field_rel_err(m.residual(), o.get("div_tconf_upts"))
    print("  mixed residual", err, flush=True)
    assert np.all(err < 2e-5), err
