"""CPU restatement of the learned LES closure's point physics and of its training
samples (test infrastructure only: imported by tests/, never by the product).

Point physics (the eddy-viscosity form of the reference's GPU build,
SGS_flux_kernel, src/cuda_kernels.cu:1466-1660, with the gradients of
vis_NS_flux, :880-960): for the conservative state U and its physical gradient
g[k][m] = dU_k/dx_m,
    v_i = U_{1+i} / rho,  vg[i][j] = dv_i/dx_j = (g[1+i][j] - v_i g[0][j]) / rho,
    Sd = sym(vg) - tr(vg)/3 I,
    WALE: nu_W = (0.5 delta)^2 (Sq:Sq)^(3/2) / ((Sd:Sd)^(5/2) + (Sq:Sq)^(5/4) + 1e-10),
          Sq = sym(vg vg) - tr(vg vg)/3 I.
Network inputs: x_{3i+j} = vg[i][j] / G with G = |vg| (Frobenius), x_10 =
log(delta^2 G / nu); standardised by the model's mean and scale; nu_t = C delta^2 G
with C = out_scale * y.

Training samples: a high-order hexahedral solution is projected, element by
element, onto the Legendre polynomials of degree <= p_lo in each direction (a
sharp modal cutoff) and evaluated at the degree-p_lo Gauss points.  With the
Favre-filtered velocity v~ = (rho v)bar / rho-bar and the SGS stress
tau_ij = (rho v_i v_j)bar - (rho v_i)bar (rho v_j)bar / rho-bar, the target is the
least-squares eddy viscosity of the deviatoric stress,
    nu_exact = -(tau^d : S~d) / (2 rho-bar S~d : S~d),
and C_exact = nu_exact / (delta^2 G) with delta = filter_ratio vol^(1/3) / (p_lo+1).
The filtered velocity gradient is the element-local derivative of the degree-p_lo
interpolant of v~.
"""
import numpy as np

from oracle import sgs_mlp as om


def point_quantities(u, g):
    """u (n, 5), g (n, 5, 3) -> rho (n), vg (n, 3, 3), Sd (n, 3, 3)."""
    u = np.asarray(u, dtype=np.float64)
    g = np.asarray(g, dtype=np.float64)
    rho = u[:, 0]
    v = u[:, 1:4] / rho[:, None]
    vg = (g[:, 1:4, :] - v[:, :, None] * g[:, 0:1, :]) / rho[:, None, None]
    tr = np.trace(vg, axis1=1, axis2=2)
    Sd = 0.5 * (vg + vg.transpose(0, 2, 1)) - (tr / 3.0)[:, None, None] * np.eye(3)
    return rho, vg, Sd


def wale_nu(vg, Sd, delta):
    g2 = np.einsum("nik,nkj->nij", vg, vg)
    Sq = 0.5 * (g2 + g2.transpose(0, 2, 1)) - (np.trace(g2, axis1=1, axis2=2) / 3.0)[:, None, None] * np.eye(3)
    num = np.einsum("nij,nij->n", Sq, Sq)
    den = np.einsum("nij,nij->n", Sd, Sd)
    return 0.25 * delta * delta * num ** 1.5 / (den ** 2.5 + num ** 1.25 + 1e-10)


def features(vg, delta, rho, mu):
    """Raw network inputs (n, 10) and G (n)."""
    G = np.sqrt(np.einsum("nij,nij->n", vg, vg))
    x = np.zeros((vg.shape[0], 10))
    ok = G > 0
    x[ok, :9] = vg[ok].reshape(-1, 9) / G[ok, None]
    x[ok, 9] = np.log(delta[ok] * delta[ok] * G[ok] * rho[ok] / mu)
    return x, G


def closure_nu(u, g, delta, mu, params=None, norm=None):
    """Eddy viscosity of the learned closure at points (u, g): the network where
    its inputs lie in the training range, WALE elsewhere (and everywhere when
    params is None).  norm = (mean, scale, lo, hi, out_scale)."""
    rho, vg, Sd = point_quantities(u, g)
    delta = np.broadcast_to(np.asarray(delta, dtype=np.float64), rho.shape)
    nu_w = wale_nu(vg, Sd, delta)
    if params is None:
        return nu_w
    mean, scale, lo, hi, out_scale = norm
    x, G = features(vg, delta, rho, mu)
    use = (G > 0) & np.all((x >= np.asarray(lo, np.float32).astype(np.float64)) &
                           (x <= np.asarray(hi, np.float32).astype(np.float64)), axis=1)
    xs = (x - np.asarray(mean, np.float32).astype(np.float64)) / np.asarray(scale, np.float32).astype(np.float64)
    xs[~use] = 0.0
    y = om.forward(params, xs.astype(np.float32), rounding=True)
    y = y.astype(np.float32).astype(np.float64)
    nu_net = y * float(np.float32(out_scale)) * delta * delta * G
    return np.where(use, nu_net, nu_w)


# ---------------------------------------------------------------------------
# training samples


def modal_filter_1d(x_hi, x_lo):
    """(n_lo, n_hi): nodal values at the points x_hi -> the truncated Legendre
    expansion (degrees < n_lo) evaluated at the points x_lo."""
    n_hi, n_lo = len(x_hi), len(x_lo)
    V_hi = np.polynomial.legendre.legvander(np.asarray(x_hi), n_hi - 1)
    V_lo = np.polynomial.legendre.legvander(np.asarray(x_lo), n_hi - 1)
    T = np.zeros((n_hi, n_hi))
    T[:n_lo, :n_lo] = np.eye(n_lo)
    return V_lo @ T @ np.linalg.inv(V_hi)


def lagrange_derivative_1d(x):
    """(n, n): D[a][b] = l_b'(x_a) of the Lagrange basis on the points x."""
    n = len(x)
    D = np.zeros((n, n))
    for a in range(n):
        for b in range(n):
            if a == b:
                D[a, b] = sum(1.0 / (x[b] - x[m]) for m in range(n) if m != b)
            else:
                num = np.prod([x[a] - x[m] for m in range(n) if m not in (a, b)])
                den = np.prod([x[b] - x[m] for m in range(n) if m != b])
                D[a, b] = num / den
    return D


def samples(u_hi, x_hi, x_lo, drdx, detj, filter_ratio, mu):
    """Training samples of a hexahedral solution.

    u_hi (n_hi^3, N, 5) nodal conservative state (point index k + n j + n^2 i for
    (r, s, t) = (x[k], x[j], x[i])); drdx (N, 3, 3) dr_d/dx_m per element; detj (N).
    Returns x (N n_lo^3, 10) raw inputs, C_exact, C_wale (N n_lo^3) and the filtered
    conservative state (n_lo^3, N, 5); sample index = lo point + n_lo^3 * element."""
    n_hi, n_lo = len(x_hi), len(x_lo)
    N = u_hi.shape[1]
    F = modal_filter_1d(x_hi, x_lo)
    D = lagrange_derivative_1d(np.asarray(x_lo))
    U = np.asarray(u_hi, dtype=np.float64).reshape(n_hi, n_hi, n_hi, N, 5)  # [i][j][k] = (t, s, r)
    rho = U[..., 0]
    m = U[..., 1:4]
    prod = np.einsum("ijkea,ijkeb->ijkeab", m, m) / rho[..., None, None]

    def filt(A):
        return np.einsum("ai,bj,ck,ijke...->abce...", F, F, F, A)

    Ub = filt(U)                       # (lo, lo, lo, N, 5)
    Pb = filt(prod)                    # (lo, lo, lo, N, 3, 3)
    rb = Ub[..., 0]
    vt = Ub[..., 1:4] / rb[..., None]
    tau = Pb - np.einsum("abcei,abcej->abceij", Ub[..., 1:4], Ub[..., 1:4]) / rb[..., None, None]
    # reference-space derivatives of v~ along r (k index), s (j), t (i)
    dr = np.einsum("cm,abmeq->abceq", D, vt)
    ds = np.einsum("bm,amceq->abceq", D, vt)
    dt = np.einsum("am,mbceq->abceq", D, vt)
    dref = np.stack([dr, ds, dt], axis=-1)              # [..., i, d] = dv_i/dr_d
    vg = np.einsum("abceid,edm->abceim", dref, drdx)    # dv_i/dx_m
    nlo3 = n_lo ** 3
    vg = vg.reshape(nlo3, N, 3, 3).transpose(1, 0, 2, 3).reshape(-1, 3, 3)
    tau = tau.reshape(nlo3, N, 3, 3).transpose(1, 0, 2, 3).reshape(-1, 3, 3)
    rbf = rb.reshape(nlo3, N).T.reshape(-1)
    tr = np.trace(vg, axis1=1, axis2=2)
    Sd = 0.5 * (vg + vg.transpose(0, 2, 1)) - (tr / 3.0)[:, None, None] * np.eye(3)
    taud = tau - (np.trace(tau, axis1=1, axis2=2) / 3.0)[:, None, None] * np.eye(3)
    ss = np.einsum("nij,nij->n", Sd, Sd)
    nu_ex = np.where(ss > 0, -np.einsum("nij,nij->n", taud, Sd) / (2.0 * rbf * np.where(ss > 0, ss, 1.0)), 0.0)
    delta = np.repeat(filter_ratio * (np.asarray(detj) * 8.0) ** (1.0 / 3.0) / n_lo, nlo3)
    x, G = features(vg, delta, rbf, mu)
    sc = delta * delta * np.where(G > 0, G, 1.0)
    c_ex = np.where(G > 0, nu_ex / sc, 0.0)
    c_w = np.where(G > 0, wale_nu(vg, Sd, delta) / sc, 0.0)
    u_lo = Ub.reshape(nlo3, N, 5)
    return x, c_ex, c_w, u_lo
