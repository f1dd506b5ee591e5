"""Integral diagnostics (CalcIntegralQuantities, src/eles.cpp:6503-6643) on the
host: the library's volume cubature rules against the reference's tables
(data/cubature_hexa.dat, data/cubature_quad.dat, kept as the data fixture
cubature_rules.npz), and the oracle's integrals against the reference binary's
history.plt values (the *_integrals_* fixtures of make_goldens.py)."""
import numpy as np
import pytest

import cases

KINDS = ["kineticenergy", "vorticity", "pressuredilatation", "straincolonproduct", "devstraincolonproduct"]


def _rules():
    with np.load(cases.GOLDEN + "/tables/cubature_rules.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def integral_goldens():
    return cases.golden_names("integrals")


def _host_solver(g, rule):
    kinds = " ".join(KINDS)
    return cases.solver_of(g, hifr_host_only=1, volume_cub_order=rule,
                           integral_quantities="%d %s" % (len(KINDS), kinds))


@pytest.mark.parametrize("ele_type,rules", [(4, range(2, 9)), (1, range(2, 13))], ids=["hex", "quad"])
def test_cubature_rules_equal_reference_tables(ele_type, rules):
    """Hexes: identical to the table.  Quads: identical weights; the points of the
    4-point rule are printed truncated in the table's last (15th) decimal, so they
    agree to 1e-15."""
    tab = _rules()
    tag = "hex" if ele_type == 4 else "quad"
    g = cases.load_golden("hex_p2_ns_integrals_2x2x2" if ele_type == 4 else "quad_p3_ns_integrals_4x3")
    for r in rules:
        s = _host_solver(g, r)
        loc = s.get("loc_volume_cubpts").reshape(-1, order="F")
        w = s.get("weight_volume_cubpts")
        want_loc = tab["%s_%d_loc" % (tag, r)].reshape(-1, order="F")
        want_w = tab["%s_%d_w" % (tag, r)]
        if ele_type == 4:
            assert np.array_equal(loc, want_loc) and np.array_equal(w, want_w), r
        else:
            assert np.abs(loc - want_loc).max() <= 1.01e-15 and np.array_equal(w, want_w), r


def test_cubature_detjac_and_basis():
    """The nodal basis at the cubature points interpolates exactly (rows sum to 1)
    and detJ at the points integrates to the mesh volume."""
    g = cases.load_golden("hex_p3_ns_integrals_3x3x3")
    s = _host_solver(g, 4)
    opp = s.get("opp_volume_cubpts").reshape(64, -1, order="F")
    assert np.allclose(opp.sum(axis=1), 1.0, atol=1e-13)
    w = s.get("weight_volume_cubpts")
    detj = s.get("vol_detjac_cubpts").reshape(64, -1, order="F")
    vol = float((w[:, None] * detj).sum())
    assert vol == pytest.approx((2 * np.pi) ** 3, rel=1e-11)  # 12-decimal weights


@pytest.mark.parametrize("name", integral_goldens())
def test_oracle_integrals_match_reference(name):
    g = cases.load_golden(name)
    o = cases.oracle_of(g)
    o.set_state(g["u0"])
    o.residual()
    tab = _rules()
    tag = "hex" if int(g["ele_type"]) == 4 else "quad"
    r = int(g["vol_cub"])
    got = o.integral_quantities(tab["%s_%d_loc" % (tag, r)], tab["%s_%d_w" % (tag, r)], list(range(5)))
    want = g["integrals"]
    assert list(g["integral_names"]) == KINDS
    # the reference prints 15 significant digits; p div v is a cancelling sum
    assert np.all(np.abs(got - want) <= 1e-12 * np.abs(want)), (got - want) / want


def test_unknown_integral_quantity_is_rejected():
    g = cases.load_golden("hex_p2_ns_integrals_2x2x2")
    with pytest.raises(RuntimeError, match="not recognized"):
        cases.solver_of(g, hifr_host_only=1, integral_quantities="1 helicity")
    with pytest.raises(RuntimeError, match="cubature rule"):
        cases.solver_of(g, hifr_host_only=1, integral_quantities="1 kineticenergy", volume_cub_order=13)
