"""The unstructured tetrahedral channel workload on the GPU: a 24 x 12 x 12 hex box
over [0,4pi] x [0,2] x [0,2pi] split into 6 tets per hex (20,736 tets), interior
vertices moved by up to 10% of the y spacing, P = 3 (20 points per tet), isothermal
no-slip walls at y = 0 and 2, periodic in x and z, parabolic Mach 0.2 profile plus
5% seeded velocity noise, Re_bulk 2800, WALE, RK45, 200 steps; FP64 and mixed.

Checks:
- FP64 residual of the initial field against the oracle (1e-12 of max|R| per
  variable) and FP64 states after 10 RK45 steps against the oracle (1e-12);
- 200 steps in FP64 and in mixed precision: the walls keep the flow bounded, and
  the mixed run stays within 1e-6 (relative L2 per conserved variable, momentum
  components against the momentum vector's norm) of the FP64 run.
dt = 1e-3 (t = 0.2); the oracle runs the 200 steps stably at 1e-3 and at 5e-4.
"""
import time

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

DT = 1e-3
STEPS = 200


def _rel_l2(a, b):
    d = (a - b).reshape(-1, a.shape[-1])
    r = b.reshape(-1, b.shape[-1])
    l2 = np.sqrt((d ** 2).sum(0) / (r ** 2).sum(0))
    mom = np.sqrt((d[:, 1:4] ** 2).sum(0) / (r[:, 1:4] ** 2).sum())
    return l2, mom


@pytest.fixture(scope="module")
def channel():
    g = cases.channel_tet_case(dt=DT, steps=STEPS)
    return g


def test_tet_channel_fp64_matches_oracle(require_gpu, channel):
    g = channel
    s = cases.solver_of(g)
    assert s.metrics_per_element()  # linear tets are affine
    o = cases.oracle_of(g, threads=16)
    u0 = cases.channel_tet_state(g, s.get("pos_upts"))
    print("tet channel: %d tets, %d points" % (u0.shape[1], u0.shape[0] * u0.shape[1]), flush=True)
    s.set_state(u0)
    o.set_state(u0)
    o.residual()
    err = cases.field_rel_err(s.residual(), o.get("div_tconf_upts"))
    print("residual field errors", err, flush=True)
    assert np.all(err < 1e-12), err
    s.set_state(u0)
    o.set_state(u0)
    s.run(10)
    o.run_steps(10)
    err = cases.field_rel_err(s.get("disu_upts"), o.get("disu_upts"))
    print("10-step state field errors", err, flush=True)
    assert np.all(err < 1e-12), err


def test_tet_channel_200_steps_fp64_and_mixed(require_gpu, channel):
    g = channel
    s64 = cases.solver_of(g)
    s32 = cases.solver_of(g, hifr_precision=1)
    u0 = cases.channel_tet_state(g, s64.get("pos_upts"))
    out = {}
    for name, s in (("fp64", s64), ("mixed", s32)):
        s.set_state(u0)
        s.run(1)  # warm-up of the kernels
        s.set_state(u0)
        t = time.perf_counter()
        s.run(STEPS)
        u = s.get("disu_upts")
        ms = (time.perf_counter() - t) * 1e3 / STEPS
        out[name] = u
        print("%s: %.3f ms per RK45 step (%d steps, incl. the final download)" % (name, ms, STEPS), flush=True)
        assert np.isfinite(u).all()
        # bounded: density within 10% of 1, speed below twice the centre-line speed
        assert np.abs(u[..., 0] - 1.0).max() < 0.1
        assert np.abs(u[..., 1:4] / u[..., :1]).max() < 3.0
    l2, mom = _rel_l2(out["mixed"], out["fp64"])
    print("mixed vs fp64 after %d steps: rel L2" % STEPS, l2, "momentum components / |momentum|", mom, flush=True)
    assert l2[0] < 1e-6 and l2[4] < 1e-6 and np.all(mom < 1e-6), (l2, mom)
