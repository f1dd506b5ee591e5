"""Restart files in the reference's format (src/output.cpp:1057-1107 writer,
src/solver.cpp:390-444 and src/eles.cpp:623-716 reader; tets
src/eles_tets.cpp:749-800), on the host (no device: hifr_host_only solvers).

The *_restart_* fixtures were made by tests/golden/make_goldens.py: the library's
writer produced the initial restart file, the reference binary read it, ran two
RK45 steps and wrote its own restart file, kept here as text.  So
- the reference reads the library's files (its trajectory from them is the
  fixture; the device reproduces it in test_gpu_parity.py's RK45 golden test);
- the library reads the reference's files (below), to the last digit;
- the library writes the reference's state back byte for byte (below).
"""
import os

import numpy as np
import pytest

import cases


def _restart_goldens():
    return [n for n in cases.golden_names() if "rest_text" in cases.load_golden(n)]


def test_restart_goldens_cover_each_element_type():
    types = {int(cases.load_golden(n)["ele_type"]) for n in _restart_goldens()}
    assert types == {1, 2, 4}


@pytest.mark.parametrize("name", _restart_goldens())
def test_reads_reference_restart(tmp_path, name):
    g = cases.load_golden(name)
    path = tmp_path / "Rest_000000002_p0000.dat"
    path.write_text(str(g["rest_text"]))
    s = cases.solver_of(g, hifr_host_only=1)
    s.read_restart(path)
    # both parse the same decimal text to the nearest double
    assert np.array_equal(s.get("disu_upts"), g["u_final"])
    assert s.time() == pytest.approx(int(g["steps"]) * float(g["dt"]), rel=1e-14)


@pytest.mark.parametrize("name", _restart_goldens())
def test_writes_reference_restart_byte_for_byte(tmp_path, name):
    g = cases.load_golden(name)
    src = tmp_path / "ref.dat"
    src.write_text(str(g["rest_text"]))
    s = cases.solver_of(g, hifr_host_only=1)
    s.read_restart(src)
    out = tmp_path / "ours.dat"
    s.write_restart(out)
    assert out.read_text() == str(g["rest_text"])


@pytest.mark.parametrize("name", ["hex_p3_ns_perturbed_3x3x3", "tet_p3_ns_perturbed_2x2x2",
                                  "quad_p3_ns_perturbed_4x3"])
def test_round_trip_keeps_15_digits(tmp_path, name):
    g = cases.load_golden(name)
    s = cases.solver_of(g, hifr_host_only=1)
    rng = np.random.default_rng(5)
    u = g["u0"] * (1.0 + 1e-3 * rng.standard_normal(g["u0"].shape))
    s.set_state(u)
    path = tmp_path / "Rest.dat"
    s.write_restart(path)
    t = cases.solver_of(g, hifr_host_only=1)
    t.read_restart(path)
    want = np.array([float("%.15g" % v) for v in u.ravel()]).reshape(u.shape)
    assert np.array_equal(t.get("disu_upts"), want)


def test_reader_rejects_other_solution_points(tmp_path):
    g = cases.load_golden("hex_p2_ns_restart_2x2x1")
    path = tmp_path / "Rest.dat"
    path.write_text(str(g["rest_text"]))
    s = cases.solver_of(g, hifr_host_only=1, order=3)
    with pytest.raises(RuntimeError, match="other solution points"):
        s.read_restart(path)
    with pytest.raises(RuntimeError, match="Could not open"):
        s.read_restart(tmp_path / "missing.dat")


def test_restart_flag_reads_the_deck_iteration(tmp_path, monkeypatch):
    """restart_flag 1 with restart_iter n reads Rest_<n>_p0000.dat from the working
    directory at set-up, as the reference does."""
    g = cases.load_golden("quad_p3_euler_restart_3x3")
    (tmp_path / "Rest_000000002_p0000.dat").write_text(str(g["rest_text"]))
    monkeypatch.chdir(tmp_path)
    s = cases.solver_of(g, hifr_host_only=1, restart_flag=1, restart_iter=2)
    assert np.array_equal(s.get("disu_upts"), g["u_final"])
