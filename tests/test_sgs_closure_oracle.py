"""Pins of the learned-closure oracle (oracle/sgs_closure.py), CPU only.

- the modal filter keeps polynomials of degree <= p_lo and removes higher
  Legendre modes; the Lagrange derivative matrix is exact to degree n-1;
- velocity gradient of a manufactured variable-density field (closed form);
- WALE vanishes in pure shear and has its closed form in pure rotation;
- the training samples: for a polynomial field whose products stay within the
  high-order space, the filtered velocity gradient, the SGS stress and the exact
  eddy viscosity equal a brute-force L2 projection by Gauss quadrature.
"""
import numpy as np

from oracle import sgs_closure as sc


def test_modal_filter_and_derivative():
    x_hi = np.polynomial.legendre.leggauss(6)[0]
    x_lo = np.polynomial.legendre.leggauss(3)[0]
    F = sc.modal_filter_1d(x_hi, x_lo)
    for deg in range(3):
        assert np.allclose(F @ x_hi ** deg, x_lo ** deg, atol=1e-13)
    for deg in range(3, 6):
        P = np.polynomial.legendre.Legendre.basis(deg)(x_hi)
        assert np.allclose(F @ P, 0.0, atol=1e-13)
    D = sc.lagrange_derivative_1d(x_lo)
    for deg in range(3):
        assert np.allclose(D @ x_lo ** deg, deg * x_lo ** max(deg - 1, 0) if deg else 0 * x_lo, atol=1e-13)


def test_velocity_gradient_closed_form():
    # rho = 1 + 0.1 x, v = (2 y, -x z, 3): U = rho v
    x, y, z = 0.3, -0.7, 0.4
    rho = 1 + 0.1 * x
    u = np.array([[rho, rho * 2 * y, -rho * x * z, 3 * rho, 50.0]])
    g = np.zeros((1, 5, 3))
    g[0, 0] = [0.1, 0, 0]
    g[0, 1] = [0.1 * 2 * y, 2 * rho, 0]
    g[0, 2] = [-0.1 * x * z - rho * z, 0, -rho * x]
    g[0, 3] = [0.3, 0, 0]
    _, vg, Sd = sc.point_quantities(u, g)
    exact = np.array([[0, 2, 0], [-z, 0, -x], [0, 0, 0]])
    assert np.allclose(vg[0], exact, atol=1e-14)
    assert abs(np.trace(Sd[0])) < 1e-14


def test_wale_shear_and_rotation():
    a, w, d = 3.0, 2.0, 0.1
    shear = np.zeros((1, 3, 3))
    shear[0, 0, 1] = a
    _, vg, Sd = sc.point_quantities(np.array([[1.0, 0, 0, 0, 1.0]]), np.zeros((1, 5, 3)))
    Sd_s = 0.5 * (shear + shear.transpose(0, 2, 1))
    assert sc.wale_nu(shear, Sd_s, np.array([d]))[0] == 0.0
    rot = np.zeros((1, 3, 3))
    rot[0, 0, 1], rot[0, 1, 0] = -w, w
    num = 2 * w ** 4 / 3
    exact = 0.25 * d * d * num ** 1.5 / (num ** 1.25 + 1e-10)
    assert np.isclose(sc.wale_nu(rot, np.zeros((1, 3, 3)), np.array([d]))[0], exact, rtol=1e-14)


def _project(f, x_lo, nq=12):
    """L2 projection of f(r, s, t) on Legendre degrees < len(x_lo) per direction,
    evaluated at the tensor points of x_lo: brute-force Gauss quadrature."""
    n = len(x_lo)
    q, wq = np.polynomial.legendre.leggauss(nq)
    R, S, T = np.meshgrid(q, q, q, indexing="ij")
    W = np.einsum("i,j,k->ijk", wq, wq, wq)
    vals = f(R, S, T)
    out = np.zeros((n, n, n))
    for a in range(n):
        for b in range(n):
            for c in range(n):
                La = np.polynomial.legendre.Legendre.basis(a)
                Lb = np.polynomial.legendre.Legendre.basis(b)
                Lc = np.polynomial.legendre.Legendre.basis(c)
                coef = np.sum(W * vals * La(R) * Lb(S) * Lc(T)) * (2 * a + 1) * (2 * b + 1) * (2 * c + 1) / 8.0
                XI, XJ, XK = np.meshgrid(x_lo, x_lo, x_lo, indexing="ij")  # [t][s][r]
                out += coef * La(XK) * Lb(XJ) * Lc(XI)
    return out  # [i=t][j=s][k=r]


def test_samples_match_brute_force_projection():
    x_hi = np.polynomial.legendre.leggauss(6)[0]
    x_lo = np.polynomial.legendre.leggauss(3)[0]
    h = 0.5  # element [0, h]^3: x = (r + 1) h / 2
    drdx = np.eye(3)[None] * (2.0 / h)
    detj = np.array([(h / 2) ** 3])
    # rho = 1 + 0.1 r, v = (s t + r, r^2 - t, 0.5 s): products rho v_i v_j have degree <= 5
    rho_f = lambda r, s, t: 1 + 0.1 * r
    v_f = [lambda r, s, t: s * t + r, lambda r, s, t: r * r - t, lambda r, s, t: 0.5 * s + 0 * r]
    T, S, R = np.meshgrid(x_hi, x_hi, x_hi, indexing="ij")
    u = np.zeros((6 ** 3, 1, 5))
    rho = rho_f(R, S, T)
    u[:, 0, 0] = rho.ravel()
    for i in range(3):
        u[:, 0, 1 + i] = (rho * v_f[i](R, S, T)).ravel()
    u[:, 0, 4] = 40.0
    mu = 1e-3
    x, c_ex, c_w, u_lo = sc.samples(u, x_hi, x_lo, drdx, detj, 1.0, mu)
    # brute force: filtered rho, momenta, products at the lo points
    rb = _project(lambda r, s, t: rho_f(r, s, t), x_lo)
    mb = [_project(lambda r, s, t, i=i: rho_f(r, s, t) * v_f[i](r, s, t), x_lo) for i in range(3)]
    assert np.allclose(u_lo[:, 0, 0], rb.ravel(), atol=1e-12)
    for i in range(3):
        assert np.allclose(u_lo[:, 0, 1 + i], mb[i].ravel(), atol=1e-12)
    tau = np.zeros((27, 3, 3))
    for i in range(3):
        for j in range(3):
            pb = _project(lambda r, s, t, i=i, j=j: rho_f(r, s, t) * v_f[i](r, s, t) * v_f[j](r, s, t), x_lo)
            tau[:, i, j] = (pb - mb[i] * mb[j] / rb).ravel()
    # filtered velocity gradient: derivative of the degree-2 interpolant of v~ at the
    # lo points, by fitting the 27 values exactly
    vt = np.stack([m.ravel() / rb.ravel() for m in mb], axis=1)
    TT, SS, RR = np.meshgrid(x_lo, x_lo, x_lo, indexing="ij")
    pts = np.stack([RR.ravel(), SS.ravel(), TT.ravel()], axis=1)
    V = np.stack([np.polynomial.legendre.legval(pts[:, 0], np.eye(3)[a]) *
                  np.polynomial.legendre.legval(pts[:, 1], np.eye(3)[b]) *
                  np.polynomial.legendre.legval(pts[:, 2], np.eye(3)[c])
                  for c in range(3) for b in range(3) for a in range(3)], axis=1)
    coef = np.linalg.solve(V, vt)
    dV = []
    for dd in range(3):
        cols = []
        for c in range(3):
            for b in range(3):
                for a in range(3):
                    idx = [a, b, c]
                    f = [np.polynomial.legendre.legval(pts[:, k], np.eye(3)[idx[k]]) for k in range(3)]
                    f[dd] = np.polynomial.legendre.legval(pts[:, dd], np.polynomial.legendre.legder(np.eye(3)[idx[dd]]))
                    cols.append(f[0] * f[1] * f[2])
        dV.append(np.stack(cols, axis=1) @ coef)  # (27, 3): d v_i / d r_dd
    vg = np.stack(dV, axis=2) * (2.0 / h)        # [p][i][m]
    Sd = 0.5 * (vg + vg.transpose(0, 2, 1)) - (np.trace(vg, axis1=1, axis2=2) / 3)[:, None, None] * np.eye(3)
    taud = tau - (np.trace(tau, axis1=1, axis2=2) / 3)[:, None, None] * np.eye(3)
    nu = -np.einsum("nij,nij->n", taud, Sd) / (2 * rb.ravel() * np.einsum("nij,nij->n", Sd, Sd))
    G = np.sqrt(np.einsum("nij,nij->n", vg, vg))
    delta = (detj[0] * 8) ** (1 / 3) / 3
    assert np.allclose(x[:, :9], vg.reshape(27, 9) / G[:, None], atol=1e-11)
    assert np.allclose(x[:, 9], np.log(delta ** 2 * G * rb.ravel() / mu), atol=1e-11)
    assert np.allclose(c_ex, nu / (delta ** 2 * G), rtol=1e-9, atol=1e-12)
    assert np.all(c_w >= 0)


# ---- the C oracle's supplied eddy viscosity (SGS_model 10, calc_sgsf_nu) ----

def _nu_oracle(L, les=1, order=3, n=4):
    import frmesh
    from oracle import pyoracle as po
    xv, conn = frmesh.box_hex(n, n, n, L, L, L)
    c2v = frmesh.gambit_to_internal(4, conn)
    deck = po.default_deck(equation=0, viscous=1, order=order, adv_type=0, dt=1e-3, tau=1.0, pen_fact=0.5,
                           vcjh_scheme=1, eta=0.0, gamma=1.4, prandtl=0.72, Mach_free_stream=0.1,
                           nx_free_stream=1.0, Re_free_stream=1600.0, L_free_stream=1.0, T_free_stream=300.0,
                           dx_cyclic=L, dy_cyclic=L, dz_cyclic=L, LES=les, SGS_model=10 if les else 0,
                           filter_ratio=1.0, fix_vis=1.0)
    return po.Oracle(deck, 4, xv, c2v)


def test_supplied_nu_zero_is_laminar():
    """nu_t = 0 adds nothing: the residual equals the laminar one bit for bit."""
    import frmesh
    L = 2 * np.pi
    lam, les = _nu_oracle(L, les=0), _nu_oracle(L, les=1)
    pos = lam.get("pos_upts")
    u0 = frmesh.tgv_state(pos, p0=1.0 / (1.4 * 0.1 ** 2))
    u0 = u0 * (1.0 + frmesh.seeded_perturbation(u0.shape, 1e-2, 3))
    lam.set_state(u0)
    les.set_state(u0)
    les.set_sgs_nu(np.zeros(pos.shape[:2]))
    assert np.array_equal(lam.residual(), les.residual())
    assert np.all(les.get("sgsf_upts") == 0.0)


def test_supplied_nu_closed_form_on_linear_fields():
    """rho, the velocity and the pressure linear in x: at P = 3 the corrected
    gradient of rho, rho v and E (degree <= 3) is exact on elements away from the
    periodic boundary, so the SGS flux there has its closed form
        F(1+i, j) = -2 rho nu_t Sd_ij,   Sd = sym(A) - tr(A)/3 I,
        F(4, j)   = -gamma rho nu_t / 0.5 * de/dx_j,  e = p / ((gamma-1) rho),
    with a different nu_t at every point; sgsf_upts holds JGinv . F."""
    L, gam = 4.0, 1.4
    o = _nu_oracle(L, les=1)
    pos = o.get("pos_upts")                       # (Nu, N, 3)
    X = pos - L / 2
    r = np.array([0.05, -0.03, 0.02])             # d rho / dx
    A = np.array([[0.3, -0.2, 0.1], [0.4, -0.1, 0.25], [-0.15, 0.2, 0.05]])   # dv_i / dx_j
    b = np.array([1.0, -0.5, 0.25])
    c = np.array([0.7, -0.4, 0.9])                # dp / dx
    rho = 1.0 + X @ r
    v = b + np.einsum("ij,...j->...i", A, X)
    p = 20.0 + X @ c
    u0 = np.concatenate([rho[..., None], rho[..., None] * v,
                         (p / (gam - 1) + 0.5 * rho * (v ** 2).sum(-1))[..., None]], axis=-1)
    nu = 1e-3 * (1.0 + np.random.default_rng(9).random(pos.shape[:2]))
    o.set_state(u0)
    o.set_sgs_nu(nu)
    o.residual()
    sg = o.get("sgsf_upts")                       # (Nu, N, 5, 3), transformed
    JG = o.get("JGinv_upts")                      # (3, 3, Nu, N)
    Sd = 0.5 * (A + A.T) - np.trace(A) / 3 * np.eye(3)
    de = (c[None, None, :] * rho[..., None] - p[..., None] * r[None, None, :]) / (rho[..., None] ** 2 * (gam - 1))
    F = np.zeros(pos.shape[:2] + (5, 3))
    F[..., 1:4, :] = -2.0 * (rho * nu)[..., None, None] * Sd
    F[..., 4, :] = -gam * (rho * nu)[..., None] / 0.5 * de
    Ft = np.einsum("lmpe,pekm->pekl", JG, F)
    # elements with no face on the periodic boundary (element centres strictly inside)
    cen = pos.mean(axis=0)
    inner = np.all((cen > L / 4) & (cen < 3 * L / 4), axis=1)
    assert inner.sum() == 8
    err = np.abs(sg[:, inner] - Ft[:, inner]).max() / np.abs(Ft[:, inner]).max()
    assert err < 1e-11, err
    # a transposed strain or a dropped density-gradient term would miss by O(1)
    assert np.abs(Ft[:, inner]).max() > 1e-4
