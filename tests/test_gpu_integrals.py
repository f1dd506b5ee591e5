"""Integral diagnostics on the device (hifr_integral_quantities; the reference's
CalcIntegralQuantities, src/eles.cpp:6503-6643, with the integral_quantities and
volume_cub_order deck keys).

- against the reference binary's history.plt values and the oracle on the same
  mesh and state (*_integrals_* fixtures), both metric storages: 1e-12 relative
  (p div v, a cancelling sum, 1e-10 relative);
- on the 32^3 P=3 Taylor-Green field at t = 0 against the closed forms: the
  kinetic energy integral pi^3, the enstrophy integral (vorticity) 3 pi^3, the
  strain-rate products 3 pi^3 (incompressible and periodic) and p div v = 0;
- the mixed-precision context against the FP64 one.
"""
import math

import numpy as np
import pytest

import cases
import frmesh
import hifr

pytestmark = pytest.mark.gpu

KINDS = ["kineticenergy", "vorticity", "pressuredilatation", "straincolonproduct", "devstraincolonproduct"]
TWO_PI = 6.2831853071795862
TOL = np.array([1e-12, 1e-12, 1e-10, 1e-12, 1e-12])


def _deck(rule):
    return dict(volume_cub_order=rule, integral_quantities="%d %s" % (len(KINDS), " ".join(KINDS)))


@pytest.mark.parametrize("metrics", [0, 1], ids=["auto", "per_point"])
@pytest.mark.parametrize("name", cases.golden_names("integrals"))
def test_integrals_match_reference(require_gpu, name, metrics):
    g = cases.load_golden(name)
    print("case", name, "metrics", metrics, flush=True)
    s = cases.solver_of(g, hifr_metrics=metrics, **_deck(int(g["vol_cub"])))
    s.set_state(g["u0"])
    got = s.integrals()
    want = g["integrals"]
    rel = np.abs(got - want) / np.abs(want)
    print("relative differences", rel)
    assert np.all(rel < TOL), rel
    # the same again (the diagnostics leave the state and the stepping untouched)
    assert np.array_equal(s.integrals(), got)
    assert np.array_equal(s.get("disu_upts"), g["u0"])


def _tgv(precision=0, n=32):
    deck = hifr.deck_text(viscous=1, order=3, vcjh_scheme_hexa=1, dt=1e-3, adv_type=3, tau=1.0, pen_fact=0.5,
                          gamma=1.4, prandtl=0.71, Mach_free_stream=0.1, nx_free_stream=1.0,
                          Re_free_stream=1600.0, L_free_stream=1.0, T_free_stream=300.0, fix_vis=1,
                          dx_cyclic=TWO_PI, dy_cyclic=TWO_PI, dz_cyclic=TWO_PI, hifr_precision=precision,
                          **_deck(5))
    s = hifr.Solver.box(deck, hifr.HEX, (n, n, n), (TWO_PI,) * 3, (0.0,) * 3)
    s.set_state(frmesh.tgv_state(s.get("pos_upts"), p0=1.0 / (1.4 * 0.1 ** 2)))
    return s


def test_tgv_32cubed_closed_forms(require_gpu):
    s = _tgv()
    got = s.integrals()
    pi3 = math.pi ** 3
    print("TGV integrals", got)
    assert got[0] == pytest.approx(pi3, rel=1e-6)
    assert got[1] == pytest.approx(3 * pi3, rel=1e-4)
    assert got[3] == pytest.approx(3 * pi3, rel=1e-4)
    assert got[4] == pytest.approx(3 * pi3, rel=1e-4)
    # p div v: p0 = 71.4 times a divergence that vanishes to discretisation error
    assert abs(got[2]) < 1e-3 * 71.4 * 3 * pi3


def test_integrals_mixed_close_to_fp64(require_gpu):
    a = _tgv(0).integrals()
    b = _tgv(1).integrals()
    rel = np.abs(a - b) / np.abs(a)
    print("mixed vs FP64", rel)
    assert np.all(rel[[0, 1, 3, 4]] < 1e-6), rel


def test_solver_executable_writes_history(require_gpu, tmp_path):
    """hifr_run on a deck with integral_quantities writes history.plt in the
    reference's column layout (HistoryOutput, src/output.cpp:1525-1616): iteration,
    log10 residual norms, forces, CL, CD, the diagnostics, physical and wall time;
    the diagnostics equal the library's after the same steps."""
    import os
    import subprocess
    g = cases.load_golden("hex_p2_ns_integrals_2x2x2")
    ele_type, xv, c2v = cases.mesh_of(g)
    groups = [("CYCLIC", frmesh.boundary_faces(ele_type, c2v))]
    frmesh.write_gmsh(str(tmp_path / "mesh.msh"), xv, c2v, ele_type, bc_groups=groups)
    deck = cases.product_deck(g, mesh_file="mesh.msh", adv_type=3, dt=1e-3, n_steps=4, monitor_res_freq=2,
                              restart_flag=0, ic_form=7, **_deck(4))
    (tmp_path / "input").write_text(deck)
    exe = os.path.join(os.path.dirname(hifr.LIB_PATH), "hifr_run")
    r = subprocess.run([exe, "input"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = (tmp_path / "history.plt").read_text().strip().split("\n")
    assert lines[0].startswith("TITLE") and lines[2].startswith("ZONE")
    header = lines[1].split(",")
    assert len(header) == 1 + 5 + 3 + 2 + len(KINDS) + 2
    rows = [[float(v) for v in ln.split(",")] for ln in lines[3:]]
    assert [int(r[0]) for r in rows] == [1, 2, 4]
    s = hifr.Solver.from_deck(str(tmp_path / "input"))
    s.run(4)
    last = np.array(rows[-1])
    assert np.allclose(last[1 + 5 + 5:1 + 5 + 5 + len(KINDS)], s.integrals(), rtol=1e-13, atol=0)
    assert last[-2] == pytest.approx(4e-3, rel=1e-12)
    assert np.allclose(last[6:11], 0.0)
