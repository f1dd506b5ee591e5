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
