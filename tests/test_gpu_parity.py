"""Parity of the gfx950 FR residual and RK update with the oracle and with the
reference's own numbers.

Tolerance (FP64 path): the residual per conserved variable within 1e-12 of
max|R| (normwise), RK states within 1e-12; reference fixtures carry only 15
printed digits, which is below these tolerances.
"""
import json
import os

import numpy as np
import pytest

import cases
import frmesh
import hifr
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _pair(g, **over):
    return cases.solver_of(g, **over), cases.oracle_of(g)


@pytest.mark.parametrize("metrics", [0, 1], ids=["auto", "per_point"])
@pytest.mark.parametrize("name", cases.golden_names("residual"))
def test_residual_matches_reference_and_oracle(require_gpu, name, metrics):
    g = cases.load_golden(name)
    s, o = _pair(g, hifr_metrics=metrics)
    print("case", name, "metrics", metrics, flush=True)
    # box meshes are affine: auto keeps one metric set per element; hex and quad
    # meshes with perturbed interior vertices have non-affine elements and keep
    # per-point metrics (linear tets are affine whatever their vertices)
    J = s.get("JGinv_upts")
    affine = np.abs(J - J[:, :, :1, :]).max() < 1e-13 * np.abs(J).max()
    if "perturbed" in name and int(g["ele_type"]) != 2:
        assert not affine
    if int(g["ele_type"]) == 2:
        assert affine
    assert s.metrics_per_element() == (metrics == 0 and affine)
    s.set_state(g["u0"])
    div = s.residual()
    J = s.get("detjac_upts")[:, :, None]
    R = div / J
    assert cases.field_rel_err(R, g["R"]).max() < TOL
    o.set_state(g["u0"])
    o.residual()
    assert cases.field_rel_err(div, o.get("div_tconf_upts")).max() < TOL
    # traces of u and of the corrected gradient (intermediate outputs)
    assert cases.field_rel_err(s.get("disu_fpts"), o.get("disu_fpts")).max() < 1e-14
    if int(g["viscous"]):
        gs, go = s.get("grad_disu_fpts"), o.get("grad_disu_fpts")
        nd = gs.shape[3]
        for d in range(nd):
            assert cases.field_rel_err(gs[..., d], go[..., d]).max() < TOL


@pytest.mark.parametrize("name", cases.golden_names("residual"))
def test_forward_euler_step_matches_reference(require_gpu, name):
    """The residual goldens come from one reference forward-Euler step with a huge
    dt (u1 = u0 - dt R); one device step of the same deck (adv_type 0, the
    advance kernels' forward-Euler branch) recovers R = (u0 - u1) / dt."""
    g = cases.load_golden(name)
    s, _ = _pair(g)
    print("case", name, "dt", float(g["dt"]), flush=True)
    s.set_state(g["u0"])
    s.run(1)
    R = (g["u0"] - s.get("disu_upts")) / float(g["dt"])
    assert cases.field_rel_err(R, g["R"]).max() < TOL


@pytest.mark.parametrize("name", cases.golden_names("rk45"))
def test_rk45_matches_reference(require_gpu, name):
    g = cases.load_golden(name)
    s, _ = _pair(g)
    s.set_state(g["u0"])
    s.run(int(g["steps"]))
    assert cases.field_rel_err(s.get("disu_upts"), g["u_final"]).max() < TOL


def _random_case(ele_type, n, order, viscous, vcjh, perturb, seed, L=cases.TWO_PI):
    if ele_type == 4:
        xv, conn = frmesh.box_hex(n[0], n[1], n[2], L, L, L)
    else:
        xv, conn = frmesh.box_quad(n[0], n[1], L, L)
    if perturb:
        xv = frmesh.perturb_interior(xv, perturb * L / max(n), seed)
    g = dict(ele_type=ele_type, xv=xv, conn_gambit=conn, order=order, vcjh=vcjh, viscous=viscous,
             tau=1.0, pen_fact=0.5, cyc=np.array([L, L, L]), dt=2e-3, kind="rk45")
    return g


@pytest.mark.parametrize("ele_type,n,order,viscous,vcjh", [
    (4, (4, 3, 5), 3, 1, 2), (4, (3, 3, 3), 2, 1, 1), (4, (3, 2, 3), 4, 1, 3), (4, (2, 2, 3), 5, 1, 2),
    (4, (3, 4, 2), 1, 1, 1), (4, (4, 3, 3), 3, 0, 3),
    (1, (5, 4), 1, 1, 1), (1, (4, 5), 2, 0, 2), (1, (6, 5), 3, 1, 2), (1, (3, 4), 4, 1, 3),
    (1, (4, 3), 5, 0, 1), (1, (3, 3), 6, 1, 2)])
def test_random_meshes_against_oracle(require_gpu, ele_type, n, order, viscous, vcjh):
    g = _random_case(ele_type, n, order, viscous, vcjh, perturb=0.1, seed=order + 10 * ele_type)
    print("case", ele_type, n, order, viscous, vcjh, flush=True)
    s, o = _pair(g)
    pos = o.get("pos_upts")
    if ele_type == 4:
        u = frmesh.tgv_state(pos, p0=1.0 / (1.4 * 0.1 ** 2))
    else:
        u = frmesh.shu_vortex_state(pos - np.pi)
    u = u * (1.0 + frmesh.seeded_perturbation(u.shape, 1e-2, 99))
    s.set_state(u)
    o.set_state(u)
    o.residual()
    assert cases.field_rel_err(s.residual(), o.get("div_tconf_upts")).max() < TOL
    # two RK45 steps
    o.set_state(u)
    o.zero_rk_register()
    o.run_steps(2)
    s.set_state(u)
    s.run(2)
    assert cases.field_rel_err(s.get("disu_upts"), o.get("disu_upts")).max() < TOL
    assert cases.field_rel_err(s.get("rk_register"), o.get("rk_register")).max() < 1e-10


@pytest.mark.parametrize("n,order,viscous", [((3, 3, 3), 3, 1), ((5, 3, 3), 3, 0), ((3, 3, 3), 4, 1)])
def test_affine_boxes_with_a_partial_last_block(require_gpu, n, order, viscous):
    """Unperturbed boxes keep one metric set per element, and at P=3 the advance
    kernels stage their flux two directions at a time in blocks of two elements;
    an odd element count leaves the last block half empty.  Residual and two RK45
    steps against the oracle (FP64), and the mixed residual within 2e-5."""
    g = _random_case(4, n, order, viscous, 1, perturb=0.0, seed=0)
    print("case affine", n, order, viscous, flush=True)
    s, o = _pair(g)
    assert s.metrics_per_element()
    pos = o.get("pos_upts")
    u = frmesh.tgv_state(pos, p0=1.0 / (1.4 * 0.1 ** 2))
    u = u * (1.0 + frmesh.seeded_perturbation(u.shape, 1e-2, 7))
    s.set_state(u)
    o.set_state(u)
    o.residual()
    R = o.get("div_tconf_upts")
    assert cases.field_rel_err(s.residual(), R).max() < TOL
    o.set_state(u)
    o.zero_rk_register()
    o.run_steps(2)
    s.set_state(u)
    s.run(2)
    assert cases.field_rel_err(s.get("disu_upts"), o.get("disu_upts")).max() < TOL
    m = cases.solver_of(g, hifr_precision=1)
    m.set_state(u)
    assert np.all(cases.field_rel_err(m.residual(), R) < 2e-5)


def test_tgv_regression_norms(require_gpu, tmp_path):
    """The reference's regression vector: Taylor-Green deck, 15^3 hexes, P=3, SD,
    25 RK45 steps (testcases/regression_tests.py:300-307), run from a deck file."""
    with open(os.path.join(cases.GOLDEN, "tgv_regression.json")) as fh:
        reg = json.load(fh)
    m, d = reg["mesh"], reg["deck"]
    L = m["length"]
    xv, conn = frmesh.box_hex(*m["n"], L, L, L, digits=m["digits"])
    c2v = frmesh.gambit_to_internal(4, conn)
    bfaces = [(e, frmesh.internal_face_to_gambit(4, f)) for e, f in frmesh.boundary_faces(4, c2v)]
    frmesh.write_gambit(str(tmp_path / "tgv.neu"), xv, conn, 4, bc_groups=[("Cyclic", bfaces)])
    deck = hifr.deck_text(viscous=1, order=d["order"], ic_form=d["ic_form"], dt=d["dt"], adv_type=3,
                          tau=d["tau"], pen_fact=d["pen_fact"], vcjh_scheme_hexa=d["vcjh_scheme_hexa"],
                          gamma=d["gamma"], prandtl=d["prandtl"], Mach_free_stream=d["Mach_free_stream"],
                          nx_free_stream=d["nx_free_stream"], Re_free_stream=d["Re_free_stream"],
                          L_free_stream=d["L_free_stream"], T_free_stream=d["T_free_stream"], fix_vis=1,
                          dx_cyclic=L, dy_cyclic=L, dz_cyclic=L, mesh_file="tgv.neu", res_norm_type=1,
                          n_steps=reg["test_iter"])
    (tmp_path / "input_tgv").write_text(deck)
    s = hifr.Solver.from_deck(str(tmp_path / "input_tgv"))
    s.run(1)
    np.testing.assert_allclose(s.norms(1), reg["step1_vals"], atol=reg["tol"], rtol=0)
    s.run(reg["test_iter"] - 1)
    np.testing.assert_allclose(s.norms(1), reg["test_vals"], atol=reg["tol"], rtol=0)


def test_deterministic(require_gpu):
    g = cases.load_golden("hex_p3_ns_perturbed_3x3x3")
    s, _ = _pair(g, adv_type=3, dt=1e-3)
    out = []
    for _ in range(2):
        s.set_state(g["u0"])
        s.run(2)
        out.append(s.get("disu_upts"))
    assert np.all(np.isfinite(out[0]))
    assert np.array_equal(out[0], out[1])


def test_entry_points_fail_loudly(require_gpu):
    import ctypes as C
    L = hifr.lib()
    # residual before the inputs exist -> HIFR_ERR_STATE, not a crash
    cfg = (C.c_byte * 256)()
    del cfg
    assert L.hifr_residual(None) == 1  # HIFR_ERR_ARG on a null context


@pytest.mark.parametrize("name", [n for n in cases.golden_names()
                                  if str(cases.load_golden(n).get("mesh_format", "neu")) == "msh"])
def test_gmsh_deck_matches_reference(require_gpu, tmp_path, name):
    """A deck naming a Gmsh file (mesh_file mesh.msh, read by the library's Gmsh
    reader) reproduces the residual the reference computed from the same file."""
    g = cases.load_golden(name)
    ele_type, xv, c2v = cases.mesh_of(g)
    bf = cases.bfaces_of(g)
    walls = [] if bf is None else [(int(e), int(f)) for e, f, _ in bf]
    groups = [("CYCLIC", [(e, f) for e, f in frmesh.boundary_faces(ele_type, c2v) if (e, f) not in walls])]
    if walls:
        groups.append(("isotherm_fix", walls))
    frmesh.write_gmsh(str(tmp_path / "mesh.msh"), xv, c2v, ele_type, bc_groups=groups)
    (tmp_path / "input").write_text(cases.product_deck(g, mesh_file="mesh.msh"))
    print("case", name, flush=True)
    s = hifr.Solver.from_deck(str(tmp_path / "input"))
    s.set_state(g["u0"])
    R = s.residual() / s.get("detjac_upts")[:, :, None]
    assert cases.field_rel_err(R, g["R"]).max() < TOL
