"""The device boundary (include/hifr.h) driven the way the reference's classes would
drive it, with no part of the library's host solver: every array comes from the
oracle's own set-up (operators opp_0/opp_3/opp_4 of eles::set_opp_*, the metrics
of set_transforms, the interface point pairs of its face matching, the wall
points of its boundary faces, the non-dimensional constants of setup_params).

hifr_create -> hifr_set_operators -> hifr_set_metrics -> hifr_set_boundaries ->
hifr_set_interfaces -> hifr_upload_state -> hifr_residual / hifr_advance ->
hifr_download, compared with the oracle at 1e-12 (include/cuda_kernels.h:29-118
is the reference interface these entry points replace).  Then the call-order and
argument errors the boundary reports instead of exiting.
"""
import ctypes as C

import numpy as np
import pytest

import cases
import hifr

pytestmark = pytest.mark.gpu

TOL = 1e-12
OK, ERR_ARG, ERR_UNSUPPORTED, ERR_STATE = 0, 1, 2, 5


class Config(C.Structure):
    """hifr_config of include/hifr.h."""
    _fields_ = [("ele_type", C.c_int), ("order", C.c_int), ("n_eles", C.c_int), ("viscous", C.c_int),
                ("riemann_solve_type", C.c_int), ("vis_riemann_solve_type", C.c_int), ("precision", C.c_int),
                ("gamma", C.c_double), ("prandtl", C.c_double), ("mu_inf", C.c_double), ("rt_inf", C.c_double),
                ("c_sth", C.c_double), ("fix_vis", C.c_double), ("tau", C.c_double), ("pen_fact", C.c_double),
                ("const_src", C.c_double), ("metric_storage", C.c_int), ("les", C.c_int),
                ("sgs_model", C.c_int), ("filter_ratio", C.c_double), ("R_ref", C.c_double),
                ("T_wall", C.c_double)]


def _abi():
    L = hifr.lib()
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
    L.hifr_create.argtypes = [C.POINTER(Config), C.POINTER(vp)]
    L.hifr_destroy.argtypes = [vp]
    L.hifr_set_operators.argtypes = [vp, dp, dp, dp]
    L.hifr_set_metrics.argtypes = [vp, dp, dp, dp, dp]
    L.hifr_set_boundaries.argtypes = [vp, C.c_int, ip, ip]
    L.hifr_set_interfaces.argtypes = [vp, C.c_int, ip, ip]
    L.hifr_upload_state.argtypes = [vp, dp]
    L.hifr_residual.argtypes = [vp]
    L.hifr_advance.argtypes = [vp, C.c_int, C.c_int, C.c_double]
    L.hifr_download.argtypes = [vp, C.c_int, dp]
    return L


def _d(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1, order="F"))
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(C.POINTER(C.c_int32))


def _config(g, o, metrics=0):
    ph = o.phys()
    c = Config()
    c.ele_type, c.order, c.n_eles = int(g["ele_type"]), int(g["order"]), o.n_eles
    c.viscous = int(g["viscous"])
    c.precision = 0
    c.gamma, c.prandtl, c.mu_inf, c.rt_inf = ph.gamma, ph.prandtl, ph.mu_inf, ph.rt_inf
    c.c_sth, c.fix_vis, c.tau, c.pen_fact = ph.c_sth, ph.fix_vis, ph.tau, ph.pen_fact
    c.const_src, c.metric_storage = ph.const_src, metrics
    c.les, c.sgs_model, c.filter_ratio = int(g.get("les", 0)), int(g.get("sgs_model", 0)), float(g.get("filter_ratio", 1.0))
    c.R_ref, c.T_wall = ph.R_ref, ph.T_wall
    return c


def _wall_points(g, o):
    """Flux points of the fixture's wall faces: face-major numbering of the element's
    flux points (fpt = face * points per face + j), + Nfp * element."""
    bf = cases.bfaces_of(g)
    if bf is None:
        return np.zeros(0, np.int32), np.zeros(0, np.int32)
    nfpf = o.nfp // (6 if int(g["ele_type"]) == 4 else 4)
    pts = [f * nfpf + j + o.nfp * e for e, f, _ in bf for j in range(nfpf)]
    types = [t for _, _, t in bf for _ in range(nfpf)]
    return np.array(pts, np.int32), np.array(types, np.int32)


def _build(L, g, o, metrics=0):
    cfg = _config(g, o, metrics)
    ctx = C.c_void_p()
    assert L.hifr_create(C.byref(cfg), C.byref(ctx)) == OK
    keep = []
    ops = [_d(o.get(n)) for n in ("opp_0", "opp_3", "opp_4")]
    keep += ops
    assert L.hifr_set_operators(ctx, *(p for _, p in ops)) == OK
    met = [_d(o.get(n)) for n in ("detjac_upts", "JGinv_upts", "tdA_fpts", "norm_fpts")]
    keep += met
    assert L.hifr_set_metrics(ctx, *(p for _, p in met)) == OK
    bp, bt = _wall_points(g, o)
    (bpa, bpp), (bta, btp) = _i(bp), _i(bt)
    assert L.hifr_set_boundaries(ctx, len(bpa), bpp, btp) == OK
    left, right = o.face_pairs()
    (la, lp), (ra, rp) = _i(left), _i(right)
    assert L.hifr_set_interfaces(ctx, len(la), lp, rp) == OK
    return ctx, keep


def _download(L, ctx, what, shape):
    out = np.zeros(int(np.prod(shape)))
    assert L.hifr_download(ctx, what, out.ctypes.data_as(C.POINTER(C.c_double))) == OK
    return out.reshape(shape, order="F")


CASES = ["hex_p3_ns_perturbed_3x3x3", "quad_p3_channel_walls_4x3", "tet_p3_channel_wale_walls_2x2x2",
         "hex_p2_les_smagorinsky_3x2x2", "hex_p3_ns_sutherland_perturbed_2x2x2"]


@pytest.mark.parametrize("metrics", [0, 1], ids=["auto", "per_point"])
@pytest.mark.parametrize("name", CASES)
def test_device_abi_with_oracle_arrays(require_gpu, name, metrics):
    L = _abi()
    g = cases.load_golden(name)
    o = cases.oracle_of(g)
    print("case", name, "metrics", metrics, flush=True)
    ctx, keep = _build(L, g, o, metrics)
    try:
        u, up = _d(g["u0"])
        assert L.hifr_upload_state(ctx, up) == OK
        assert L.hifr_residual(ctx) == OK
        shp = (o.nu, o.n_eles, o.nf)
        div = _download(L, ctx, 2, shp)
        o.set_state(g["u0"])
        o.residual()
        assert cases.field_rel_err(div, o.get("div_tconf_upts")).max() < TOL
        # one RK45 stage through the same entry points
        assert L.hifr_upload_state(ctx, up) == OK
        assert L.hifr_advance(ctx, 0, 3, 1e-3) == OK
        o2 = cases.oracle_of(g, adv_type=3, dt=1e-3)
        o2.set_state(g["u0"])
        o2.zero_rk_register()
        o2.residual()
        o2.advance(0, 3)
        assert cases.field_rel_err(_download(L, ctx, 0, shp), o2.get("disu_upts")).max() < TOL
    finally:
        L.hifr_destroy(ctx)


def test_device_abi_call_order_and_argument_errors(require_gpu):
    L = _abi()
    g = cases.load_golden("hex_p3_ns_perturbed_3x3x3")
    o = cases.oracle_of(g)
    cfg = _config(g, o)
    ctx = C.c_void_p()
    assert L.hifr_create(C.byref(cfg), C.byref(ctx)) == OK
    try:
        u, up = _d(g["u0"])
        # nothing set yet
        assert L.hifr_residual(ctx) == ERR_STATE
        assert L.hifr_upload_state(ctx, up) == ERR_STATE
        left, right = o.face_pairs()
        (la, lp), (ra, rp) = _i(left), _i(right)
        assert L.hifr_set_interfaces(ctx, len(la), lp, rp) == ERR_STATE  # before the metrics
        ops = [_d(o.get(n)) for n in ("opp_0", "opp_3", "opp_4")]
        bad = ops[0][0].copy()
        bad[3] = np.nan
        assert L.hifr_set_operators(ctx, bad.ctypes.data_as(C.POINTER(C.c_double)), ops[1][1], ops[2][1]) != OK
        assert L.hifr_set_operators(ctx, *(p for _, p in ops)) == OK
        met = [_d(o.get(n)) for n in ("detjac_upts", "JGinv_upts", "tdA_fpts", "norm_fpts")]
        assert L.hifr_set_metrics(ctx, *(p for _, p in met)) == OK
        # interface tables: an index out of range, a point used twice, a point left over
        big = la.copy()
        big[0] = o.nfp * o.n_eles
        assert L.hifr_set_interfaces(ctx, len(la), big.ctypes.data_as(C.POINTER(C.c_int32)), rp) == ERR_ARG
        dup = ra.copy()
        dup[1] = dup[0]
        assert L.hifr_set_interfaces(ctx, len(la), lp, dup.ctypes.data_as(C.POINTER(C.c_int32))) == ERR_ARG
        assert L.hifr_set_interfaces(ctx, len(la) - 1, lp, rp) == ERR_UNSUPPORTED
        assert L.hifr_residual(ctx) == ERR_STATE
        assert L.hifr_set_interfaces(ctx, len(la), lp, rp) == OK
        # only isothermal walls are built as boundary conditions
        p1, pp1 = _i([0])
        t1, tp1 = _i([7])
        assert L.hifr_set_boundaries(ctx, 1, pp1, tp1) == ERR_UNSUPPORTED
        assert L.hifr_upload_state(ctx, up) == OK
        assert L.hifr_residual(ctx) == OK
        assert L.hifr_advance(ctx, 0, 1, 1e-3) == ERR_UNSUPPORTED  # adv_type 1 is not a time integrator here
        assert L.hifr_download(ctx, 99, up) == ERR_ARG
        assert L.hifr_download(ctx, 0, None) == ERR_ARG
    finally:
        L.hifr_destroy(ctx)
    # configurations that are not built
    for field, value in (("ele_type", 3), ("order", 9), ("vis_riemann_solve_type", 1), ("precision", 7)):
        c = _config(g, o)
        setattr(c, field, value)
        assert L.hifr_create(C.byref(c), C.byref(ctx)) == ERR_UNSUPPORTED, field
    c = _config(g, o)
    c.n_eles = 0
    assert L.hifr_create(C.byref(c), C.byref(ctx)) == ERR_ARG
