"""hifr_run replays every time step but the last as a captured HIP graph (one graph
per trace-buffer parity, for one adv_type and dt). Each replay must give the state
of the eager, stage-by-stage path bit for bit: replays several times in a row and
between downloads, eager steps in between, after a new state (which drops the
graphs: the mixed path's reference state is a kernel argument), after a change of
dt, and with forward Euler, whose single stage flips the trace parity every step.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

CASES = [("hex_p3_ns_perturbed_3x3x3", {}), ("hex_p3_les_wale_perturbed_3x3x3", {}),
         ("tet_p3_channel_wale_walls_2x2x2", {}), ("quad_p3_ns_perturbed_4x3", {}),
         ("quad_p3_ns_roe_channel_walls_4x3", {}), ("hex_p3_ns_tgv_2x2x2", {}),
         ("hex_p3_ns_perturbed_3x3x3", {"hifr_precision": 1}), ("tet_p3_ns_perturbed_2x2x2", {"hifr_precision": 1})]


def _state(g, scale=1.0):
    u = g["u0"].copy()
    if scale != 1.0:
        u[..., 1:-1] *= scale
    return u


def _eager(s, u0, n, nst):
    """States after 1..n eager steps (every stage through hifr_advance)."""
    s.set_state(u0)
    out = []
    for _ in range(n):
        for st in range(nst):
            s.advance(st)
        out.append(s.get("disu_upts").copy())
    return out


@pytest.mark.parametrize("adv_type", [3, 0], ids=["rk45", "euler"])
@pytest.mark.parametrize("name,over", CASES, ids=["%s%s" % (n, "_mixed" if o else "") for n, o in CASES])
def test_graph_replays_equal_eager_steps(require_gpu, name, over, adv_type):
    g = cases.load_golden(name)
    nst = 5 if adv_type == 3 else 1
    print("case", name, over, "adv_type", adv_type, flush=True)
    s = cases.solver_of(g, adv_type=adv_type, dt=1e-4, **over)
    u0 = _state(g)
    ref = _eager(s, u0, 9, nst)

    s.set_state(u0)
    s.run(2)  # one replay, then the eager last step
    assert np.array_equal(s.get("disu_upts"), ref[1])
    s.run(3)  # replays of both parities' graphs
    assert np.array_equal(s.get("disu_upts"), ref[4])
    for st in range(nst):  # an eager step between replays
        s.advance(st)
    assert np.array_equal(s.get("disu_upts"), ref[5])
    s.run(3)
    assert np.array_equal(s.get("disu_upts"), ref[8])

    # a new state drops the graphs (the mixed reference state changes with it)
    u1 = _state(g, 1.25)
    ref1 = _eager(s, u1, 4, nst)
    s.set_state(u1)
    s.run(4)
    assert np.array_equal(s.get("disu_upts"), ref1[3])


@pytest.mark.parametrize("with_net", [False, True], ids=["wale_fallback", "network"])
def test_graph_replays_learned_closure(require_gpu, with_net):
    """The learned closure's kernels in the captured step: the gradient-trace
    kernel stores each point's eddy viscosity and the flux kernel of the same
    stage reads it back; with a network attached (which drops the graphs) its
    weights are read inside the graph."""
    import frmesh
    import hifr
    from oracle import sgs_mlp as om
    TWO_PI = cases.TWO_PI
    xv, conn = frmesh.box_hex(3, 3, 3, TWO_PI, TWO_PI, TWO_PI)
    g = dict(ele_type=4, xv=xv, conn_gambit=conn, order=3, vcjh=1, viscous=1, tau=1.0, pen_fact=0.5,
             cyc=np.array([TWO_PI] * 3), dt=1e-4, kind="residual", les=1, sgs_model=1, filter_ratio=2.0,
             Re=5000.0)
    s = cases.solver_of(g, adv_type=3, dt=1e-4, hifr_learned_sgs=1)
    pos = s.get("pos_upts")
    u0 = frmesh.tgv_state(pos, p0=1.0 / (1.4 * 0.1 ** 2))
    u0 = u0 * (1.0 + frmesh.seeded_perturbation(u0.shape, 2e-2, 5))
    if with_net:
        m = hifr.MLP()
        p = om.init_params(3)
        p[om.N_PARAMS - 1] = -3.0
        m.set_params(p)
        mean, scale = np.zeros(10, np.float32), np.ones(10, np.float32)
        mean[9], scale[9] = 5.0, 2.0
        m.set_normalization(mean, scale, np.full(10, -1e30, np.float32), np.full(10, 1e30, np.float32), 0.1)
        s.run(3)  # graphs captured without the network, then dropped by the attach
        s.set_sgs_network(m)
    print("learned closure, network" if with_net else "learned closure, WALE fallback", flush=True)
    ref = _eager(s, u0, 6, 5)
    s.set_state(u0)
    s.run(4)
    assert np.array_equal(s.get("disu_upts"), ref[3])
    s.run(2)
    assert np.array_equal(s.get("disu_upts"), ref[5])


def test_graph_follows_a_change_of_dt(require_gpu):
    """hifr_run with one dt, then another on the same context (no new state in
    between): the second run captures new graphs instead of replaying the first's,
    and the state equals hifr_advance with the same two dts."""
    import ctypes as C
    import hifr
    L = hifr.lib()
    L.hifr_advance.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double]
    g = cases.load_golden("hex_p3_ns_perturbed_3x3x3")
    u0 = _state(g)
    e = cases.solver_of(g, adv_type=3, dt=1e-4)
    e.set_state(u0)
    for dt in (1e-4, 2e-4):
        for _ in range(3):
            for st in range(5):
                assert L.hifr_advance(C.c_void_p(e.ctx()), st, 3, C.c_double(dt)) == 0
    r = cases.solver_of(g, adv_type=3, dt=1e-4)
    r.set_state(u0)
    for dt in (1e-4, 2e-4):
        assert L.hifr_run(C.c_void_p(r.ctx()), 3, 3, C.c_double(dt)) == 0
    assert np.array_equal(r.get("disu_upts"), e.get("disu_upts"))
