"""Mathematical pins of the oracle, independent of its own formulas.

Each check is chosen so that a plausible mistake fails it:
  * Gauss table          -- numpy's Gauss-Legendre nodes.
  * opp_0 / opp_4        -- exact interpolation / differentiation of degree-P polynomials.
  * opp_3, DG (eta=0)    -- the DG-FR identity g_L'(x_j) = -l_j(-1)/w_j with Gauss weights.
  * opp_3, SD            -- the closed form g_L = (-1)^P/2 (1-x) P_P(x) (eta = P/(P+1)).
  * opp_3, any eta       -- sum_j w_j g_L'(x_j) = g_L(1) - g_L(-1) = -1.
  * metrics              -- total volume = box volume on perturbed meshes; closed surfaces
                            (sum of n tdA = 0, sum of x.n tdA = d * volume).
  * residual             -- free-stream preservation, discrete conservation, manufactured
                            solutions for the Euler flux (advected vortex) and for the viscous
                            stress / heat flux terms (shear layer, conduction), with the
                            expected order of convergence.
"""
import math

import numpy as np
import pytest

import cases
import frmesh
from oracle import pyoracle as po

TWO_PI = cases.TWO_PI


def _periodic_box(ele_type, n, L=TWO_PI, perturb=0.0, seed=0):
    if ele_type == 4:
        xv, conn = frmesh.box_hex(n, n, n, L, L, L)
    else:
        xv, conn = frmesh.box_quad(n, n, L, L)
    if perturb:
        xv = frmesh.perturb_interior(xv, perturb * L / n, seed)
    return xv, frmesh.gambit_to_internal(ele_type, conn)


def _oracle(ele_type, n, order, viscous=0, vcjh=1, perturb=0.0, L=TWO_PI, **kw):
    xv, c2v = _periodic_box(ele_type, n, L, perturb)
    d = dict(equation=0, viscous=viscous, order=order, adv_type=0, dt=1e-3, tau=kw.pop("tau", 1.0),
             pen_fact=0.5, vcjh_scheme=vcjh, eta=0.0, dx_cyclic=L, dy_cyclic=L, dz_cyclic=L)
    d.update(cases.PHYS)
    d.update(kw)
    return po.Oracle(po.default_deck(**d), ele_type, xv, c2v)


def _weights_1d(order):
    return np.polynomial.legendre.leggauss(order + 1)[1]


@pytest.mark.parametrize("order", range(0, 7))
def test_gauss_table_matches_legendre_roots(order):
    o = _oracle(1, 2, order)
    x = o.get("loc_upts")[0, : order + 1]
    ref = np.polynomial.legendre.leggauss(order + 1)[0]
    assert np.abs(x - ref).max() <= 1.2e-15


@pytest.mark.parametrize("ele_type,order", [(1, 1), (1, 3), (1, 5), (4, 2), (4, 3), (4, 4)])
def test_interpolation_and_derivative_exact(ele_type, order):
    o = _oracle(ele_type, 2, order)
    nd = 3 if ele_type == 4 else 2
    rng = np.random.default_rng(order)
    coef = rng.standard_normal((nd, order + 1))
    P = np.polynomial.polynomial

    def f(loc):
        return np.prod([P.polyval(loc[d], coef[d]) for d in range(nd)], axis=0)

    def df(loc, d):
        parts = [P.polyval(loc[k], coef[k]) if k != d else P.polyval(loc[k], P.polyder(coef[k]))
                 for k in range(nd)]
        return np.prod(parts, axis=0)

    lu, lf = o.get("loc_upts"), o.get("tloc_fpts")
    u = f(lu)
    np.testing.assert_allclose(o.get("opp_0") @ u, f(lf), rtol=0, atol=1e-12 * np.abs(u).max())
    D = o.get("opp_4")
    for d in range(nd):
        np.testing.assert_allclose(D[:, :, d] @ u, df(lu, d), rtol=0, atol=1e-11 * np.abs(u).max())


def _left_end_column(o, order):
    """opp_3 column of the quad flux point at r = -1 on the line s = x_0 (face 3, last point)."""
    n = order + 1
    fpL = 3 * n + order
    return o.get("opp_3")[:n, fpL], o.get("opp_0")[fpL, :n]


@pytest.mark.parametrize("order", range(1, 6))
def test_dg_correction_is_the_dg_lift(order):
    o = _oracle(1, 2, order, vcjh=1)
    col, lm1 = _left_end_column(o, order)
    # column = -g_L'(x_j) * l_0(x_0) = -g_L'(x_j);  DG: g_L'(x_j) = -l_j(-1) / w_j
    np.testing.assert_allclose(col, lm1 / _weights_1d(order), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("order", range(1, 6))
def test_sd_correction_closed_form(order):
    o = _oracle(1, 2, order, vcjh=2)
    col, _ = _left_end_column(o, order)
    x = o.get("loc_upts")[0, : order + 1]
    L = np.polynomial.legendre
    cP = np.zeros(order + 1)
    cP[order] = 1.0
    gLd = (-1) ** order / 2.0 * (-L.legval(x, cP) + (1 - x) * L.legval(x, L.legder(cP)))
    np.testing.assert_allclose(-col, gLd, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("vcjh", [1, 2, 3])
@pytest.mark.parametrize("order", [2, 3, 4])
def test_correction_integrates_to_minus_one(vcjh, order):
    o = _oracle(1, 2, order, vcjh=vcjh)
    col, _ = _left_end_column(o, order)
    assert abs(np.dot(_weights_1d(order), -col) + 1.0) < 1e-13


@pytest.mark.parametrize("ele_type", [1, 4])
def test_metrics_volume_and_closed_surfaces(ele_type):
    n, order = 3, 3
    o = _oracle(ele_type, n, order, perturb=0.12)
    nd = 3 if ele_type == 4 else 2
    w1 = _weights_1d(order)
    wu = np.einsum("i,j,k->kji", w1, w1, w1).ravel() if nd == 3 else np.outer(w1, w1).ravel()
    vol = np.einsum("j,je->e", wu, o.get("detjac_upts"))
    assert abs(vol.sum() - TWO_PI ** nd) < 1e-11 * TWO_PI ** nd
    assert vol.min() > 0
    wf = np.outer(w1, w1).ravel() if nd == 3 else w1
    nfaces = 2 * nd
    wfp = np.tile(wf, nfaces)
    tdA, nrm, pos = o.get("tdA_fpts"), o.get("norm_fpts"), o.get("pos_fpts")
    flux = np.einsum("f,fe,fed->ed", wfp, tdA, nrm)
    assert np.abs(flux).max() < 1e-12
    xn = np.einsum("f,fe,fed,fed->e", wfp, tdA, nrm, pos)
    np.testing.assert_allclose(xn, nd * vol, rtol=1e-12)


def _state_const(o, rho=1.2, vel=(0.3, -0.2, 0.1), p=2.0, gamma=1.4):
    nd = o.nd
    u = np.zeros((o.nu, o.n_eles, o.nf))
    u[..., 0] = rho
    for i in range(nd):
        u[..., 1 + i] = rho * vel[i]
    u[..., nd + 1] = p / (gamma - 1) + 0.5 * rho * sum(v * v for v in vel[:nd])
    return u


@pytest.mark.parametrize("ele_type,viscous", [(1, 0), (1, 1), (4, 0), (4, 1)])
def test_free_stream_preservation(ele_type, viscous):
    o = _oracle(ele_type, 3, 3, viscous=viscous, perturb=0.12, vcjh=2)
    o.set_state(_state_const(o))
    div = o.residual()
    assert np.abs(div).max() < 1e-12 * 10.0


@pytest.mark.parametrize("ele_type,viscous", [(1, 1), (4, 1)])
def test_discrete_conservation(ele_type, viscous):
    o = _oracle(ele_type, 3, 3, viscous=viscous, perturb=0.1, vcjh=3)
    pos = o.get("pos_upts")
    u = _state_const(o)
    rng = np.random.default_rng(3)
    u *= 1.0 + 0.05 * np.sin(pos[..., :1] + 2 * pos[..., 1:2]) + 0.01 * rng.uniform(-1, 1, u.shape)
    o.set_state(u)
    div = o.residual()
    w1 = _weights_1d(3)
    wu = np.einsum("i,j,k->kji", w1, w1, w1).ravel() if o.nd == 3 else np.outer(w1, w1).ravel()
    total = np.einsum("j,jek->k", wu, div)
    scale = np.einsum("j,jek->k", wu, np.abs(div))
    assert np.all(np.abs(total) < 1e-13 * scale + 1e-14)


def _vortex_error(n, order):
    L = 20.0
    xv, conn = frmesh.box_quad(n, n, L, L, origin=(-10.0, -10.0))
    c2v = frmesh.gambit_to_internal(1, conn)
    d = po.default_deck(equation=0, viscous=0, order=order, adv_type=0, dt=1e-3, vcjh_scheme=1,
                        dx_cyclic=L, dy_cyclic=L)
    o = po.Oracle(d, 1, xv, c2v)
    pos = o.get("pos_upts")
    o.set_state(frmesh.shu_vortex_state(pos))
    R = o.residual() / o.get("detjac_upts")[:, :, None]
    h = 1e-5
    dudt = (frmesh.shu_vortex_state(pos, t=h) - frmesh.shu_vortex_state(pos, t=-h)) / (2 * h)
    e = R + dudt
    return math.sqrt((e ** 2).mean() / (dudt ** 2).mean())


def test_euler_residual_converges_for_advected_vortex():
    # relative L2 truncation error of the residual, order P once the core (radius 1)
    # is resolved: 40^2 -> 80^2 elements on [-10, 10]^2
    e1, e2 = _vortex_error(40, 3), _vortex_error(80, 3)
    order = math.log(e1 / e2, 2)
    assert e2 < 5e-3 and order > 2.8, (e1, e2, order)


def _viscous_field(o, case):
    """State and exact viscous contribution to div/J of two fields:
    'shear': rho=1, u = U sin(y), p const  -> only tau_xy and viscous heating;
    'heat':  u = 0, p const, rho = 1/(1 + eps sin x) -> only heat conduction."""
    ph = o.phys()
    mu, gam, Pr = ph.mu_inf, ph.gamma, ph.prandtl
    pos = o.get("pos_upts")
    x, y = pos[..., 0], pos[..., 1]
    nd = o.nd
    u = np.zeros((o.nu, o.n_eles, o.nf))
    exact = np.zeros_like(u)
    p = 50.0
    if case == "shear":
        U = 2.0
        vel = U * np.sin(y)
        u[..., 0] = 1.0
        u[..., 1] = vel
        u[..., nd + 1] = p / (gam - 1) + 0.5 * vel ** 2
        exact[..., 1] = mu * U * np.sin(y)
        exact[..., nd + 1] = -mu * U ** 2 * np.cos(2 * y)
    else:
        eps = 0.1
        u[..., 0] = 1.0 / (1.0 + eps * np.sin(x))
        u[..., nd + 1] = p / (gam - 1)
        exact[..., nd + 1] = (mu * gam / Pr) * p * eps * np.sin(x) / (gam - 1)
    return u, exact


def _manufactured_viscous(ele_type, n, case):
    """Viscous part of the residual (Navier-Stokes minus Euler on the same state)
    against the exact viscous divergence; returns (relative error of the active
    component, max |viscous part of the density residual|)."""
    ons = _oracle(ele_type, n, 3, viscous=1, vcjh=1, tau=0.0)
    oeu = _oracle(ele_type, n, 3, viscous=0, vcjh=1)
    u, exact = _viscous_field(ons, case)
    ons.set_state(u)
    oeu.set_state(u)
    J = ons.get("detjac_upts")[:, :, None]
    Rv = ons.residual() / J - oeu.residual() / J
    k = 1 if case == "shear" else ons.nd + 1
    err = math.sqrt(((Rv[..., k] - exact[..., k]) ** 2).mean() / (exact[..., k] ** 2).mean())
    return err, np.abs(Rv[..., 0]).max()


@pytest.mark.parametrize("ele_type", [1, 4])
@pytest.mark.parametrize("case", ["shear", "heat"])
def test_viscous_terms_manufactured(ele_type, case):
    # second-derivative terms: the pointwise residual error of the viscous part is
    # O(h^(P-1)) at P = 3; a wrong sign, factor or dropped term leaves an O(1) error
    e1, r1 = _manufactured_viscous(ele_type, 4, case)
    e2, r2 = _manufactured_viscous(ele_type, 8, case)
    assert e2 < 0.08 and e1 / e2 > 2.5, (e1, e2)
    assert r2 == 0.0


def _channel_oracle(ele_type, order, T_wall, perturb=0.1):
    """Channel [0, 4pi] x [0, 2] (x [0, 2pi]) with isothermal walls at y = 0, 2."""
    if ele_type == 4:
        xv, conn = frmesh.box_hex(3, 2, 2, 2 * TWO_PI, 2.0, TWO_PI)
        cyc = dict(dx_cyclic=2 * TWO_PI, dy_cyclic=float("inf"), dz_cyclic=TWO_PI)
    else:
        xv, conn = frmesh.box_quad(4, 3, 2 * TWO_PI, 2.0)
        cyc = dict(dx_cyclic=2 * TWO_PI, dy_cyclic=float("inf"), dz_cyclic=float("inf"))
    xv = frmesh.perturb_interior(xv, perturb * 0.5, seed=5)
    c2v = frmesh.gambit_to_internal(ele_type, conn)
    walls = frmesh.wall_faces(ele_type, c2v, xv, 1, (0.0, 2.0))
    d = dict(equation=0, viscous=1, order=order, adv_type=0, dt=1e-3, tau=1.0, pen_fact=0.5, vcjh_scheme=1,
             eta=0.0, **cyc)
    d.update(cases.PHYS)
    d.update(Mach_free_stream=0.2, T_wall=T_wall)
    return po.Oracle(po.default_deck(**d), ele_type, xv, c2v, bfaces=[(e, f, 11) for e, f in walls])


def _quiescent(o, T):
    """Gas at rest, rho = 1, temperature T (non-dimensional): p = rho R_ref T."""
    ph = o.phys()
    u = np.zeros((o.nu, o.n_eles, o.nf))
    u[..., 0] = 1.0
    u[..., o.nf - 1] = ph.R_ref * T / (ph.gamma - 1.0)
    return u


@pytest.mark.parametrize("ele_type", [1, 4])
def test_wall_keeps_isothermal_gas_at_rest(ele_type):
    # gas at rest at the wall temperature: ghost state = trace, no heat flux,
    # so the residual vanishes to round-off (on a perturbed mesh)
    o = _channel_oracle(ele_type, 3, T_wall=300.0)
    u = _quiescent(o, o.phys().T_wall)
    o.set_state(u)
    o.residual()
    R = o.get("div_tconf_upts")
    assert np.abs(R).max() < 1e-10 * np.abs(u).max(), np.abs(R).max()


@pytest.mark.parametrize("ele_type", [1, 4])
def test_hot_wall_heats_the_gas(ele_type):
    # walls 10% hotter than the gas at rest (T_wall = 1.1): heat flows in
    # (dE/dt = -div/J > 0 next to the walls), the momentum stays at rest (equal
    # pressure on both sides of the wall), and the LDG penalty tau (u_l - u_r) on
    # the density -- the ghost density p/(R_ref T_wall) is lower -- draws mass out
    # of the wall cells, as the reference's Dirichlet LDG flux does for every field
    o = _channel_oracle(ele_type, 3, T_wall=330.0)
    u = _quiescent(o, 1.0)
    o.set_state(u)
    o.residual()
    dudt = -o.get("div_tconf_upts") / o.get("detjac_upts")[:, :, None]
    y = o.get("pos_upts")[..., 1]
    near = (y < 0.3) | (y > 1.7)
    e = dudt[..., o.nf - 1]
    assert e[near].mean() > 0 and e[near].mean() > 10 * abs(e[~near].mean())
    assert np.abs(dudt[..., 1:o.nf - 1]).max() < 1e-10 * np.abs(e).max()
    assert dudt[..., 0][near].mean() < 0


# --------------------------------------------------------------------------
# tetrahedra: operators pinned by exact polynomial calculus on the reference
# simplex (r, s, t >= -1, r + s + t <= -1), independent of the Dubiner formulas

def _tet_oracle(order, perturb=0.0, viscous=1):
    xv, hexes = frmesh.box_hex(2, 2, 2, TWO_PI, TWO_PI, TWO_PI)
    if perturb:
        xv = frmesh.perturb_interior(xv, perturb * TWO_PI / 2, seed=9)
    tets = frmesh.hex_to_tets(xv, hexes)
    d = dict(equation=0, viscous=viscous, order=order, adv_type=0, dt=1e-3, tau=1.0, pen_fact=0.5,
             vcjh_scheme=1, eta=0.0, dx_cyclic=TWO_PI, dy_cyclic=TWO_PI, dz_cyclic=TWO_PI)
    d.update(cases.PHYS)
    return po.Oracle(po.default_deck(**d), 2, xv, tets)


def _monomials(P):
    return [(a, b, c) for a in range(P + 1) for b in range(P + 1 - a) for c in range(P + 1 - a - b)]


def _mono(x, e):
    return x[0] ** e[0] * x[1] ** e[1] * x[2] ** e[2]


def _tet_rule(m=8):
    """Collapsed Gauss rule on the reference tet, exact to degree 2m-3."""
    g, w = np.polynomial.legendre.leggauss(m)
    pts, wts = [], []
    for i in range(m):
        for j in range(m):
            for k in range(m):
                a, b, c = g[i], g[j], g[k]
                r = 0.25 * (1 + a) * (1 - b) * (1 - c) - 1
                s = 0.5 * (1 + b) * (1 - c) - 1
                t = c
                pts.append((r, s, t))
                wts.append(w[i] * w[j] * w[k] * 0.125 * (1 - b) * (1 - c) ** 2)
    return np.array(pts), np.array(wts)


def _tri_rule(m=8):
    g, w = np.polynomial.legendre.leggauss(m)
    pts, wts = [], []
    for i in range(m):
        for j in range(m):
            a, b = g[i], g[j]
            pts.append((0.5 * (1 + a) * (1 - b) - 1, b))
            wts.append(w[i] * w[j] * 0.5 * (1 - b))
    return np.array(pts), np.array(wts)


def _nodal_basis_eval(upts, P, x):
    """Values at x (n,3) of the Lagrange basis on the solution points, built from
    monomials (an independent construction of the reference's nodal basis)."""
    E = _monomials(P)
    V = np.array([[_mono(p, e) for e in E] for p in upts])
    M = np.array([[_mono(p, e) for e in E] for p in x])
    return M @ np.linalg.inv(V)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_tet_interpolation_and_derivative_exact(order):
    o = _tet_oracle(order)
    up = o.get("loc_upts").T
    fp = o.get("tloc_fpts").T
    I, D = o.get("opp_0"), o.get("opp_4")
    for e in _monomials(order):
        v = np.array([_mono(p, e) for p in up])
        assert np.allclose(I @ v, [_mono(p, e) for p in fp], atol=1e-12)
        for c in range(3):
            de = list(e)
            if e[c] == 0:
                exact = np.zeros(len(up))
            else:
                de[c] -= 1
                exact = e[c] * np.array([_mono(p, de) for p in up])
            assert np.allclose(D[..., c] @ v, exact, atol=1e-11)


@pytest.mark.parametrize("order", [1, 2, 3])
def test_tet_dg_correction_is_the_face_lift(order):
    # DG (c_tet = 0): the correction field L_f of flux point f is the L2
    # projection onto P_P(tet) of its face Lagrange polynomial, i.e. for every v in
    # P_P: int_tet v L_f = int_face v l_f dA.  The reference's face rule has
    # 15-digit weights (sum 2 - 2e-13), so the identity holds to ~1e-12.
    o = _tet_oracle(order)
    up = o.get("loc_upts").T
    fp = o.get("tloc_fpts").T
    L = o.get("opp_3")  # (nu, nfp)
    nfpf = (order + 1) * (order + 2) // 2
    q3, w3 = _tet_rule()
    B = _nodal_basis_eval(up, order, q3)  # (nq, nu)
    q2, w2 = _tri_rule()
    faces = [(lambda a, b: (a, -1 - a - b, b), np.sqrt(3.0)), (lambda a, b: (-1.0, b, a), 1.0),
             (lambda a, b: (a, -1.0, b), 1.0), (lambda a, b: (b, a, -1.0), 1.0)]
    face_coords = [lambda p: (p[0], p[2]), lambda p: (p[2], p[1]), lambda p: (p[0], p[2]), lambda p: (p[1], p[0])]
    E2 = [(a, b) for a in range(order + 1) for b in range(order + 1 - a)]
    worst = 0.0
    for f in range(4):
        fx = np.array([face_coords[f](p) for p in fp[f * nfpf:(f + 1) * nfpf]])
        V2 = np.array([[x[0] ** a * x[1] ** b for a, b in E2] for x in fx])
        C2 = np.linalg.inv(V2)  # face Lagrange coefficients in monomials
        M2 = np.array([[x[0] ** a * x[1] ** b for a, b in E2] for x in q2]) @ C2  # (nq2, nfpf)
        x3 = np.array([faces[f][0](a, b) for a, b in q2])
        for e in _monomials(order):
            vq = np.array([_mono(p, e) for p in q3])
            lhs = (w3 * vq) @ (B @ L[:, f * nfpf:(f + 1) * nfpf])
            rhs = faces[f][1] * ((w2 * np.array([_mono(p, e) for p in x3])) @ M2)
            worst = max(worst, np.abs(lhs - rhs).max())
    assert worst < 5e-12, worst


@pytest.mark.parametrize("order", [2, 3])
def test_tet_free_stream_and_conservation(order):
    o = _tet_oracle(order, perturb=0.08)
    u = np.zeros((o.nu, o.n_eles, 5))
    u[..., 0], u[..., 1], u[..., 2], u[..., 3] = 1.0, 0.3, -0.2, 0.1
    u[..., 4] = 71.4 / 0.4 + 0.5 * 0.14
    o.set_state(u)
    o.residual()
    assert np.abs(o.get("div_tconf_upts")).max() < 1e-9
    u = frmesh.tgv_state(o.get("pos_upts"), p0=1 / (1.4 * 0.01))
    o.set_state(u)
    o.residual()
    R = o.get("div_tconf_upts")
    # integral of the reference-space divergence over each element with the
    # weights W_j = int l_j (exact for the degree-P interpolant) sums to zero
    q3, w3 = _tet_rule()
    W = w3 @ _nodal_basis_eval(o.get("loc_upts").T, order, q3)
    total = np.einsum("j,jek->k", W, R)
    scale = np.einsum("j,jek->k", W, np.abs(R))
    assert np.all(np.abs(total) < 1e-12 * scale), (total, scale)


def test_set_metrics_replaces_the_computed_metrics():
    """or_set_metrics: given its own metrics back, the oracle's residual is
    unchanged bit for bit; given per-element metrics of an affine mesh (each
    element's first-point values everywhere), it changes only by rounding."""
    import cases as _cases
    g = _cases.load_golden("hex_p3_ns_tgv_2x2x2")
    o = _cases.oracle_of(g)
    o.set_state(g["u0"])
    R0 = o.residual().copy()
    m = {k: o.get(k) for k in ("detjac_upts", "JGinv_upts", "tdA_fpts", "norm_fpts")}
    o.set_metrics(m["detjac_upts"], m["JGinv_upts"], m["tdA_fpts"], m["norm_fpts"])
    assert np.array_equal(o.residual(), R0)
    nu = m["detjac_upts"].shape[0]
    dj = np.repeat(m["detjac_upts"][:1], nu, axis=0)
    jg = np.repeat(m["JGinv_upts"][:, :, :1], nu, axis=2)
    o.set_metrics(dj, jg, m["tdA_fpts"], m["norm_fpts"])
    R1 = o.residual()
    err = _cases.field_rel_err(R1, R0)
    assert err.max() < 1e-11 and not np.array_equal(R1, R0), err


# ---- Roe flux (riemann_solve_type 2) ----
# The 2-D Roe flux is pinned by the reference-binary fixtures (quad_*_roe_*: the
# residual of Euler and Navier-Stokes decks with walls, and an RK45 trajectory).
# The reference has no 3-D Roe; the oracle's 3-D form is pinned below by what the
# mathematics fixes: it reduces to the 2-D flux when the third velocity and normal
# components vanish, it commutes with rotations of the frame, and it is consistent.


def _roe_states(rng, nd, n=64):
    rho = rng.uniform(0.5, 2.0, (n, 2))
    v = rng.uniform(-0.8, 0.8, (n, 2, nd))
    p = rng.uniform(0.5, 3.0, (n, 2))
    E = p / 0.4 + 0.5 * rho * (v ** 2).sum(-1)
    u = np.concatenate([rho[..., None], rho[..., None] * v, E[..., None]], axis=-1)
    nrm = rng.normal(size=(n, nd))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    return u, nrm


def _euler_flux_n(u, n, gamma=1.4):
    nd = len(n)
    v = u[1:1 + nd] / u[0]
    p = (gamma - 1.0) * (u[nd + 1] - 0.5 * u[0] * (v @ v))
    vn = v @ n
    return np.concatenate([[u[0] * vn], u[1:1 + nd] * vn + p * n, [(u[nd + 1] + p) * vn]])


def test_roe_3d_reduces_to_2d():
    rng = np.random.default_rng(71)
    u2, n2 = _roe_states(rng, 2)
    for (ul, ur), n in zip(u2, n2):
        f2 = po.common_flux(2, ul, ur, n)
        lift = lambda u: np.array([u[0], u[1], u[2], 0.0, u[3]])
        f3 = po.common_flux(2, lift(ul), lift(ur), np.array([n[0], n[1], 0.0]))
        assert np.array_equal(f3[[0, 1, 2, 4]], f2) and f3[3] == 0.0


def test_roe_3d_rotation_invariant():
    """F(Q u_L, Q u_R, Q n) = Q F(u_L, u_R, n) for a rotation Q acting on the momentum:
    a swapped or dropped velocity component, or a wrong normal index, breaks it."""
    rng = np.random.default_rng(72)
    u3, n3 = _roe_states(rng, 3)
    for (ul, ur), n in zip(u3, n3):
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        rot = lambda u: np.concatenate([[u[0]], Q @ u[1:4], [u[4]]])
        fa = po.common_flux(2, rot(ul), rot(ur), Q @ n)
        fb = rot(po.common_flux(2, ul, ur, n))
        assert np.allclose(fa, fb, rtol=0, atol=1e-12 * np.abs(fb).max())


def test_roe_3d_from_planar_2d_in_every_plane():
    """A 2-D state embedded in the (y, z) or (x, z) plane gives the 2-D flux in that
    plane's components (the 2-D flux itself is pinned by the reference)."""
    rng = np.random.default_rng(73)
    u2, n2 = _roe_states(rng, 2)
    for plane in ((1, 2), (0, 2), (2, 0)):
        for (ul, ur), n in zip(u2, n2):
            def emb(u):
                out = np.zeros(5)
                out[0], out[4] = u[0], u[3]
                out[1 + plane[0]], out[1 + plane[1]] = u[1], u[2]
                return out
            n3 = np.zeros(3)
            n3[plane[0]], n3[plane[1]] = n
            f3 = po.common_flux(2, emb(ul), emb(ur), n3)
            f2 = po.common_flux(2, ul, ur, n)
            want = emb(f2)
            assert np.allclose(f3, want, rtol=0, atol=1e-13 * np.abs(f2).max())


@pytest.mark.parametrize("nd", [2, 3])
def test_roe_consistent_and_upwind(nd):
    """F(u, u, n) = f(u).n; for a supersonic state moving along n the Roe flux is
    the left state's flux (all waves run to the right)."""
    rng = np.random.default_rng(74 + nd)
    u, nrm = _roe_states(rng, nd)
    for (ul, _), n in zip(u, nrm):
        f = po.common_flux(2, ul, ul, n)
        assert np.allclose(f, _euler_flux_n(ul, n), rtol=0, atol=1e-13 * np.abs(f).max())
    n = np.zeros(nd)
    n[0] = 1.0
    for (ul, ur), _ in zip(u, nrm):
        for s in (ul, ur):
            s[1:1 + nd] = s[0] * np.r_[4.0, np.zeros(nd - 1)] + 0.01 * s[1:1 + nd]
            p = 1.0
            s[nd + 1] = p / 0.4 + 0.5 * (s[1:1 + nd] @ s[1:1 + nd]) / s[0]
        f = po.common_flux(2, ul, ur, n)
        assert np.allclose(f, _euler_flux_n(ul, n), rtol=1e-12, atol=1e-12)


def test_roe_goldens_pin_the_flux():
    """The Roe fixtures are far from what the Rusanov flux gives, so the
    reference comparisons above pin the oracle's 2-D Roe flux."""
    names = [n for n in cases.golden_names() if int(cases.load_golden(n).get("riemann", 0)) == 2]
    assert len(names) >= 3
    for name in names:
        g = cases.load_golden(name)
        ele_type, xv, c2v = cases.mesh_of(g)
        o = po.Oracle(cases.oracle_deck(g, riemann_solve_type=0), ele_type, xv, c2v, bfaces=cases.bfaces_of(g))
        o.set_state(g["u0"])
        if str(g["kind"]) == "residual":
            o.residual()
            R = o.get("div_tconf_upts") / o.get("detjac_upts")[:, :, None]
            assert cases.field_rel_err(R, g["R"]).max() > 1e-5
        else:
            o.zero_rk_register()
            o.run_steps(int(g["steps"]))
            assert cases.field_rel_err(o.get("disu_upts"), g["u_final"]).max() > 1e-9
