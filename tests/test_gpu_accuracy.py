"""Order of accuracy of the GPU path on the 2-D isentropic (Shu) vortex.

The vortex of the quad workload (beta = 5, unit core radius, advected at
Mach 0.5 along +x on the periodic square [-10, 10]^2) is run to t = 1 with
RK45 at dt = 0.0025 (time error far below the spatial error) on successively
refined meshes, and the L2 error of every conserved variable against the
analytically advected vortex is measured by Gauss quadrature.  FR with the DG
correction is expected to converge at order P+1 (the reference's verification,
doc/AIAApaper/aiaa_atlanta_2014/summaryTable_ele2_err1.tex, reports P+1 for the
energy field).  A dropped or mis-signed term anywhere in the residual stalls the
convergence at O(1) or first order.
"""
import numpy as np
import pytest

import frmesh
import hifr

pytestmark = pytest.mark.gpu


def _vortex_l2_error(nel, order, t_end=1.0, dt=0.0025):
    L = 20.0
    xv, conn = frmesh.box_quad(nel, nel, L, L, origin=(-10.0, -10.0))
    c2v = frmesh.gambit_to_internal(1, conn)
    deck = hifr.deck_text(viscous=0, order=order, vcjh_scheme_quad=1, dt=dt, adv_type=3, dx_cyclic=L, dy_cyclic=L)
    s = hifr.Solver.from_arrays(deck, 1, xv, c2v)
    pos = s.get("pos_upts")
    s.set_state(frmesh.shu_vortex_state(pos))
    n_steps = int(round(t_end / dt))
    s.run(n_steps)
    err = s.get("disu_upts") - frmesh.shu_vortex_state(pos, t=n_steps * dt)
    w = np.polynomial.legendre.leggauss(order + 1)[1]
    wu = np.einsum("i,j->ji", w, w).ravel()
    J = s.get("detjac_upts")
    return np.sqrt(np.einsum("j,je,jek->k", wu, J, err ** 2))


@pytest.mark.parametrize("order,meshes", [(3, (20, 40, 80)), (4, (10, 20, 40))])
def test_vortex_converges_at_order_p_plus_1(require_gpu, order, meshes):
    errs = [_vortex_l2_error(n, order) for n in meshes]
    rates = [np.log2(a / b) for a, b in zip(errs, errs[1:])]
    print("order", order, "errors", errs, "rates", rates)
    # finest pair: every variable converges at P+1 (within the pre-asymptotic spread)
    assert np.all(rates[-1] > order + 0.7), rates
    assert np.all(rates[0] > order - 0.5), rates
    # the oracle gives 2.4e-5 (P=3, 80^2) and 4.1e-5 (P=4, 40^2) on these meshes
    assert errs[-1].max() < 1e-4, errs[-1]
