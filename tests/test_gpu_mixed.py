"""Mixed-precision path (HIFR_MIXED) against the FP64 path and the oracle.

The mixed kernels keep the state, the RK register, detJ, the face metrics and
the RK update in FP64 and do everything else in FP32 on the deviation of the
state from a constant reference at rest.  Its error is measured, per conserved
variable, as the relative L2 norm ||u_mixed - u_fp64|| / ||u_fp64|| (and the
max norm likewise), against the FP64 GPU path, which itself matches the oracle
to 1e-12.  Target: <= 1e-6 per variable after the 500-step 48^3 P=4 run.
"""
import os

import numpy as np
import pytest

import cases
import frmesh
import hifr
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

TWO_PI = cases.TWO_PI


def _deck(order, precision, n_steps=1, dt=1e-3):
    return hifr.deck_text(viscous=1, order=order, vcjh_scheme_hexa=1, dt=dt, adv_type=3, tau=1.0, pen_fact=0.5,
                          gamma=1.4, prandtl=0.71, Mach_free_stream=0.1, nx_free_stream=1.0,
                          Re_free_stream=1600.0, L_free_stream=1.0, T_free_stream=300.0, fix_vis=1,
                          dx_cyclic=TWO_PI, dy_cyclic=TWO_PI, dz_cyclic=TWO_PI, hifr_precision=precision,
                          n_steps=n_steps)


def _tgv_state(pos, seed=1234):
    u = frmesh.tgv_state(pos, p0=1.0 / (1.4 * 0.1 ** 2))
    return u * (1.0 + frmesh.seeded_perturbation(u.shape, 1e-3, seed))


def _rel_errors(a, b):
    """(relative L2, relative max) per conserved variable of a against b."""
    d = (a - b).reshape(-1, a.shape[-1])
    r = b.reshape(-1, b.shape[-1])
    l2 = np.sqrt((d ** 2).sum(0) / (r ** 2).sum(0))
    mx = np.abs(d).max(0) / np.abs(r).max(0)
    return l2, mx


def _momentum_errors(a, b):
    """Relative L2 error of each momentum component against the L2 norm of the
    momentum vector: the Taylor-Green w starts at zero, so its own norm is not a
    scale for its error early in the run."""
    d = (a - b).reshape(-1, a.shape[-1])[:, 1:4]
    r = b.reshape(-1, b.shape[-1])[:, 1:4]
    return np.sqrt((d ** 2).sum(0) / (r ** 2).sum())


def _kinetic_energy(u, detj, w3):
    """Volume integral of rho |v|^2 / 2 by the Gauss rule of the solution points."""
    ke = 0.5 * (u[..., 1] ** 2 + u[..., 2] ** 2 + u[..., 3] ** 2) / u[..., 0]
    return float(np.einsum("j,je,je->", w3, detj, ke))


@pytest.mark.parametrize("name", ["hex_p3_ns_tgv_2x2x2", "hex_p3_ns_perturbed_3x3x3", "hex_p4_ns_2x2x2",
                                  "hex_p3_les_wale_perturbed_3x3x3", "hex_p3_channel_walls_3x3x2",
                                  "tet_p3_ns_perturbed_2x2x2", "tet_p3_channel_wale_walls_2x2x2"])
def test_mixed_residual_close_to_oracle(require_gpu, name):
    g = cases.load_golden(name)
    s = cases.solver_of(g, hifr_precision=1)
    o = cases.oracle_of(g)
    s.set_state(g["u0"])
    o.set_state(g["u0"])
    o.residual()
    err = cases.field_rel_err(s.residual(), o.get("div_tconf_upts"))
    print(name, "mixed residual field errors", err)
    # FP32 arithmetic on O(1) deviations: ~1e-6 of max|R| per variable
    assert np.all(err < 2e-5), err


def test_mixed_tgv_32cubed_100_steps(require_gpu):
    n, order, steps = 32, 3, 100
    s64 = hifr.Solver.box(_deck(order, 0), hifr.HEX, (n,) * 3, (TWO_PI,) * 3, (0.0,) * 3)
    s32 = hifr.Solver.box(_deck(order, 1), hifr.HEX, (n,) * 3, (TWO_PI,) * 3, (0.0,) * 3)
    u0 = _tgv_state(s64.get("pos_upts"))
    s64.set_state(u0)
    s32.set_state(u0)
    s64.run(steps)
    s32.run(steps)
    u32, u64 = s32.get("disu_upts"), s64.get("disu_upts")
    l2, mx = _rel_errors(u32, u64)
    mom = _momentum_errors(u32, u64)
    print("32^3 P=3 100 steps: rel L2", l2, "rel max", mx, "momentum components / |momentum|", mom)
    assert l2[0] < 1e-6 and l2[4] < 1e-6 and np.all(mom < 1e-6), (l2, mom)


def test_mixed_tgv_48cubed_p4_500_steps(require_gpu):
    """The mixed-precision workload: 48^3 hexes, P=4, 500 RK45 steps; per-variable
    errors against FP64 and the kinetic-energy dissipation rate.  dt = 5e-4
    (t = 0.25): the 32^3 P=3 step of 1e-3 is beyond the RK45 stability limit at
    h = 2pi/48, P = 4 (both paths diverge; the oracle shows the same limit at
    12^3, P=4: dt = 4e-3 diverges, 2e-3 is stable)."""
    n, order, steps, every, dt = 48, 4, 500, 50, 5e-4
    s64 = hifr.Solver.box(_deck(order, 0, dt=dt), hifr.HEX, (n,) * 3, (TWO_PI,) * 3, (0.0,) * 3)
    s32 = hifr.Solver.box(_deck(order, 1, dt=dt), hifr.HEX, (n,) * 3, (TWO_PI,) * 3, (0.0,) * 3)
    u0 = frmesh.tgv_state(s64.get("pos_upts"), p0=1.0 / (1.4 * 0.1 ** 2))
    s64.set_state(u0)
    s32.set_state(u0)
    detj = s64.get("detjac_upts")
    w1 = np.polynomial.legendre.leggauss(order + 1)[1]
    w3 = np.einsum("i,j,k->kji", w1, w1, w1).ravel()
    vol = TWO_PI ** 3
    ke64, ke32 = [_kinetic_energy(u0, detj, w3) / vol], [_kinetic_energy(u0, detj, w3) / vol]
    for _ in range(steps // every):
        s64.run(every)
        s32.run(every)
        ke64.append(_kinetic_energy(s64.get("disu_upts"), detj, w3) / vol)
        ke32.append(_kinetic_energy(s32.get("disu_upts"), detj, w3) / vol)
    u32, u64 = s32.get("disu_upts"), s64.get("disu_upts")
    l2, mx = _rel_errors(u32, u64)
    mom = _momentum_errors(u32, u64)
    t = np.arange(len(ke64)) * every * dt
    eps64 = -np.gradient(ke64, t)
    eps32 = -np.gradient(ke32, t)
    print("48^3 P=4 500 steps: rel L2", l2, "rel max", mx, "momentum components / |momentum|", mom)
    print("t", t.tolist())
    print("KE fp64", ke64)
    print("KE mixed", ke32)
    print("-dKE/dt fp64", eps64.tolist())
    print("-dKE/dt mixed", eps32.tolist())
    out = os.environ.get("HIFR_MIXED_REPORT")
    if out:
        import json
        with open(out, "w") as f:
            json.dump({"rel_l2": l2.tolist(), "rel_max": mx.tolist(), "momentum_rel_l2": mom.tolist(),
                       "t": t.tolist(), "ke_fp64": ke64,
                       "ke_mixed": ke32, "eps_fp64": eps64.tolist(), "eps_mixed": eps32.tolist()}, f)
    assert l2[0] < 1e-6 and l2[4] < 1e-6 and np.all(mom < 1e-6), (l2, mom)
    # the TGV at Re 1600 starts at KE = 1/8 (rho ~ 1) and dissipates
    assert abs(ke64[0] - 0.125) < 1e-3 and ke64[-1] < ke64[0]
    assert np.abs(np.array(ke32) - np.array(ke64)).max() < 1e-6 * ke64[0]
