"""Python binding (ctypes) of libhifr.so: the gfx950 FR residual / RK library.

`Solver` wraps the host solver API of include/hifr_solver.h (deck + mesh ->
elements, interfaces, device context); `ctx()` exposes the device ABI of
include/hifr.h for timing.  Arrays are returned as Fortran-ordered numpy arrays
with the reference's index order, e.g. u[j, i, k] = disu_upts(j, i, k).

The library must be built (make -C hifiles-solver_amd); there is no fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhifr.so")
_LIB = None

HEX, QUAD, TET = 4, 1, 2

_GET = {"disu_upts": 0, "div_tconf_upts": 1, "disu_fpts": 2, "grad_disu_fpts": 3, "detjac_upts": 4,
        "pos_upts": 5, "opp_0": 6, "opp_3": 7, "opp_4": 8, "JGinv_upts": 9, "tdA_fpts": 10,
        "norm_fpts": 11, "xv": 12, "c2v": 13, "rk_register": 14, "loc_volume_cubpts": 15,
        "weight_volume_cubpts": 16, "opp_volume_cubpts": 17, "vol_detjac_cubpts": 18}


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libhifr.so is not built: make -C hifiles-solver_amd")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip, cp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_char_p
        L.hifr_solver_from_deck.argtypes = [cp, C.POINTER(vp), cp, C.c_int]
        L.hifr_solver_from_arrays.argtypes = [cp, C.c_int, C.c_int, dp, C.c_int, ip, C.POINTER(vp), cp, C.c_int]
        L.hifr_solver_from_arrays_bc.argtypes = [cp, C.c_int, C.c_int, dp, C.c_int, ip, C.c_int, ip,
                                                 C.POINTER(vp), cp, C.c_int]
        L.hifr_solver_box.argtypes = [cp, C.c_int, ip, dp, dp, C.c_int, C.POINTER(vp), cp, C.c_int]
        L.hifr_solver_destroy.argtypes = [vp]
        L.hifr_solver_sizes.argtypes = [vp, ip]
        L.hifr_solver_get.argtypes = [vp, C.c_int, dp]
        L.hifr_solver_get.restype = C.c_long
        L.hifr_solver_set_state.argtypes = [vp, dp]
        L.hifr_solver_interfaces.argtypes = [vp, ip, ip]
        L.hifr_solver_interfaces.restype = C.c_long
        for f in ("hifr_solver_residual",):
            getattr(L, f).argtypes = [vp]
        L.hifr_solver_advance.argtypes = [vp, C.c_int]
        L.hifr_solver_run.argtypes = [vp, C.c_int]
        L.hifr_solver_norms.argtypes = [vp, C.c_int, dp]
        L.hifr_solver_write_restart.argtypes = [vp, cp]
        L.hifr_solver_read_restart.argtypes = [vp, cp, cp, C.c_int]
        L.hifr_solver_time.argtypes = [vp]
        L.hifr_solver_time.restype = C.c_double
        L.hifr_solver_n_integrals.argtypes = [vp]
        L.hifr_solver_integrals.argtypes = [vp, dp]
        L.hifr_solver_ctx.argtypes = [vp]
        L.hifr_solver_ctx.restype = vp
        L.hifr_solver_dt.argtypes = [vp]
        L.hifr_solver_dt.restype = C.c_double
        L.hifr_solver_adv_type.argtypes = [vp]
        L.hifr_run_timed.argtypes = [vp, C.c_int, C.c_int, C.c_double, dp, dp, ip]
        L.hifr_synchronize.argtypes = [vp]
        L.hifr_run.argtypes = [vp, C.c_int, C.c_int, C.c_double]
        L.hifr_download.argtypes = [vp, C.c_int, dp]
        L.hifr_status_string.argtypes = [C.c_int]
        L.hifr_status_string.restype = C.c_char_p
        fp = C.POINTER(C.c_float)
        L.hifr_mlp_create.argtypes = [C.POINTER(vp), C.c_ulonglong]
        L.hifr_mlp_destroy.argtypes = [vp]
        L.hifr_mlp_set_params.argtypes = [vp, fp]
        L.hifr_mlp_get_params.argtypes = [vp, fp]
        L.hifr_mlp_set_normalization.argtypes = [vp, fp, fp, fp, fp, C.c_float]
        L.hifr_mlp_loss_grad.argtypes = [vp, C.c_long, fp, fp, fp, fp]
        L.hifr_mlp_set_data.argtypes = [vp, C.c_long, fp, fp]
        L.hifr_mlp_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, fp]
        L.hifr_mlp_predict.argtypes = [vp, C.c_long, fp, fp]
        L.hifr_set_sgs_network.argtypes = [vp, vp]
        L.hifr_sgs_samples.argtypes = [vp, C.c_int, dp, dp, C.c_double, fp, fp, fp, dp]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def deck_text(**kw):
    """Input deck text (reference key names) with the keys the solver requires."""
    base = {"equation": 0, "viscous": 0, "order": 3, "mesh_file": "none.neu", "n_steps": 1,
            "riemann_solve_type": 0, "vis_riemann_solve_type": 0, "adv_type": 3, "dt_type": 0,
            "dt": 1e-3, "ic_form": 1}
    base.update(kw)
    lines = []
    for k, v in base.items():
        if isinstance(v, float):
            lines.append("%s %.17g" % (k, v))
        else:
            lines.append("%s %s" % (k, v))
    return "\n".join(lines) + "\n"


class Solver:
    """Host solver + device context for one mesh of one element type."""

    def __init__(self, handle):
        self._h = handle
        sz = (C.c_int * 6)()
        lib().hifr_solver_sizes(self._h, sz)
        self.nu, self.nfp, self.nf, self.nd, self.n_eles, self.n_inters = list(sz)

    @staticmethod
    def _check(rc, err):
        if rc != 0:
            raise RuntimeError(err.value.decode(errors="replace"))

    @classmethod
    def from_deck(cls, path):
        h, err = C.c_void_p(), C.create_string_buffer(1024)
        cls._check(lib().hifr_solver_from_deck(path.encode(), C.byref(h), err, 1024), err)
        return cls(h)

    @classmethod
    def from_arrays(cls, deck, ele_type, xv, c2v, bfaces=None):
        """bfaces: optional (n, 3) (element, local face, bctype) physical boundary faces."""
        xv = np.ascontiguousarray(xv, dtype=np.float64)
        c2v = np.ascontiguousarray(c2v, dtype=np.int32)
        h, err = C.c_void_p(), C.create_string_buffer(1024)
        ip = C.POINTER(C.c_int)
        if bfaces is not None and len(bfaces):
            bf = np.ascontiguousarray(np.asarray(bfaces, dtype=np.int32).reshape(-1, 3))
            rc = lib().hifr_solver_from_arrays_bc(deck.encode(), int(ele_type), xv.shape[0], _dp(xv), c2v.shape[0],
                                                  c2v.ctypes.data_as(ip), bf.shape[0], bf.ctypes.data_as(ip),
                                                  C.byref(h), err, 1024)
            cls._check(rc, err)
            return cls(h)
        rc = lib().hifr_solver_from_arrays(deck.encode(), int(ele_type), xv.shape[0], _dp(xv), c2v.shape[0],
                                           c2v.ctypes.data_as(ip), C.byref(h), err, 1024)
        cls._check(rc, err)
        return cls(h)

    @classmethod
    def box(cls, deck, ele_type, n, length, origin, digits=0):
        nd = 3 if ele_type == HEX else 2
        ni = (C.c_int * 3)(*(list(n) + [1] * (3 - len(n))))
        ll = (C.c_double * 3)(*(list(length) + [1.0] * (3 - len(length))))
        oo = (C.c_double * 3)(*(list(origin) + [0.0] * (3 - len(origin))))
        h, err = C.c_void_p(), C.create_string_buffer(1024)
        rc = lib().hifr_solver_box(deck.encode(), int(ele_type), ni, ll, oo, int(digits), C.byref(h), err, 1024)
        cls._check(rc, err)
        del nd
        return cls(h)

    def close(self):
        if self._h:
            lib().hifr_solver_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape_of(self, name):
        nu, nfp, N, nf, nd = self.nu, self.nfp, self.n_eles, self.nf, self.nd
        return {"disu_upts": (nu, N, nf), "div_tconf_upts": (nu, N, nf), "rk_register": (nu, N, nf),
                "disu_fpts": (nfp, N, nf), "grad_disu_fpts": (nfp, N, nf, nd), "detjac_upts": (nu, N),
                "pos_upts": (nu, N, nd), "opp_0": (nfp, nu), "opp_3": (nu, nfp), "opp_4": (nu, nu, nd),
                "JGinv_upts": (nd, nd, nu, N), "tdA_fpts": (nfp, N), "norm_fpts": (nfp, N, nd)}.get(name)

    def get(self, name):
        n = lib().hifr_solver_get(self._h, _GET[name], None)
        if n < 0:
            raise RuntimeError("hifr_solver_get(%s) failed" % name)
        out = np.zeros(n, dtype=np.float64)
        if lib().hifr_solver_get(self._h, _GET[name], _dp(out)) != n:
            raise RuntimeError("hifr_solver_get(%s) failed" % name)
        shp = self.shape_of(name)
        return out.reshape(shp, order="F") if shp else out

    def set_state(self, u):
        flat = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1, order="F"))
        if flat.size != self.nu * self.n_eles * self.nf:
            raise ValueError("state size")
        if lib().hifr_solver_set_state(self._h, _dp(flat)) != 0:
            raise RuntimeError("set_state failed")

    def interfaces(self):
        n = lib().hifr_solver_interfaces(self._h, None, None)
        left = np.zeros(n, dtype=np.int32)
        right = np.zeros(n, dtype=np.int32)
        lib().hifr_solver_interfaces(self._h, left.ctypes.data_as(C.POINTER(C.c_int)),
                                     right.ctypes.data_as(C.POINTER(C.c_int)))
        return left, right

    def residual(self):
        if lib().hifr_solver_residual(self._h) != 0:
            raise RuntimeError("residual failed")
        return self.get("div_tconf_upts")

    def advance(self, stage):
        if lib().hifr_solver_advance(self._h, int(stage)) != 0:
            raise RuntimeError("advance failed")

    def run(self, n_steps):
        if lib().hifr_solver_run(self._h, int(n_steps)) != 0:
            raise RuntimeError("run failed")

    def norms(self, norm_type=1):
        out = np.zeros(5)
        rc = lib().hifr_solver_norms(self._h, int(norm_type), _dp(out))
        if rc != 0:
            raise RuntimeError("norms failed (%d)" % rc)
        return out[: self.nf]

    def ctx(self):
        return lib().hifr_solver_ctx(self._h)

    def write_restart(self, path):
        """The state in the reference's restart format (15 significant digits)."""
        if lib().hifr_solver_write_restart(self._h, str(path).encode()) != 0:
            raise RuntimeError("write_restart failed")

    def read_restart(self, path):
        """Reads a restart file of the reference's format into the state."""
        err = C.create_string_buffer(1024)
        self._check(lib().hifr_solver_read_restart(self._h, str(path).encode(), err, 1024), err)

    def time(self):
        return lib().hifr_solver_time(self._h)

    def integrals(self):
        """The deck's integral_quantities for the current state."""
        n = lib().hifr_solver_n_integrals(self._h)
        out = np.zeros(max(n, 1))
        rc = lib().hifr_solver_integrals(self._h, _dp(out))
        if rc != 0:
            raise RuntimeError("integrals failed (%d)" % rc)
        return out[:n]

    def metrics_per_element(self):
        """True when the device keeps one metric set per element (all elements affine)."""
        flag = C.c_int(-1)
        rc = lib().hifr_metrics_per_element(C.c_void_p(self.ctx()), C.byref(flag))
        if rc != 0:
            raise RuntimeError("hifr_metrics_per_element failed (%d)" % rc)
        return bool(flag.value)

    def device_metrics(self):
        """The metrics the kernels read (hifr_download HIFR_DETJAC_UPTS ..
        HIFR_NORM_FPTS), expanded to the per-point layouts: dict of detjac_upts,
        JGinv_upts, tdA_fpts, norm_fpts."""
        out = {}
        for what, name in ((5, "detjac_upts"), (6, "JGinv_upts"), (7, "tdA_fpts"), (8, "norm_fpts")):
            shp = self.shape_of(name)
            a = np.zeros(int(np.prod(shp)), dtype=np.float64)
            rc = lib().hifr_download(C.c_void_p(self.ctx()), what, _dp(a))
            if rc != 0:
                raise RuntimeError("hifr_download(%s): %s" % (name, lib().hifr_status_string(rc).decode()))
            out[name] = a.reshape(shp, order="F")
        return out

    def dt(self):
        return lib().hifr_solver_dt(self._h)

    def adv_type(self):
        return lib().hifr_solver_adv_type(self._h)

    def run_timed(self, n_steps):
        """(total_ms, kernel_ms[3], launches[3]) for n_steps, eager launches; HIP
        events bracket each kernel of every step's first stage."""
        tot = C.c_double()
        km = (C.c_double * 3)()
        kl = (C.c_int * 3)()
        rc = lib().hifr_run_timed(self.ctx(), int(n_steps), self.adv_type(), self.dt(), C.byref(tot), km, kl)
        if rc != 0:
            raise RuntimeError("hifr_run_timed: %s" % lib().hifr_status_string(rc).decode())
        return tot.value, list(km), list(kl)

    def synchronize(self):
        lib().hifr_synchronize(self.ctx())

    def set_sgs_network(self, mlp):
        """Attach a learned eddy-viscosity network (MLP) to a deck with
        hifr_learned_sgs 1; None: the WALE fallback everywhere."""
        rc = lib().hifr_set_sgs_network(self.ctx(), mlp.handle if mlp is not None else None)
        if rc != 0:
            raise RuntimeError("hifr_set_sgs_network: " + lib().hifr_status_string(rc).decode())
        # the context copies the weights: the network may be trained further or
        # closed afterwards without affecting it

    def sgs_samples(self, order_lo, filter_ratio=1.0, want_state=False):
        """Training samples of the learned closure from the current state, filtered
        onto degree order_lo: (x (n, 10) raw inputs, C_exact (n), C_wale (n)[,
        filtered state (Nu_lo, N, 5)])."""
        nl3 = (order_lo + 1) ** 3
        n = nl3 * self.n_eles
        order = int(round(self.nu ** (1.0 / self.nd))) - 1  # hexahedra: Nu = (P+1)^3
        F = np.ascontiguousarray(modal_filter_1d(order, order_lo))
        D = np.ascontiguousarray(lagrange_derivative_1d(gauss_points(order_lo)))
        x = np.zeros((n, 10), dtype=np.float32)
        ce = np.zeros(n, dtype=np.float32)
        cw = np.zeros(n, dtype=np.float32)
        ulo = np.zeros(nl3 * self.n_eles * 5) if want_state else None
        rc = lib().hifr_sgs_samples(self.ctx(), int(order_lo), _dp(F), _dp(D), float(filter_ratio), _fp(x), _fp(ce),
                                    _fp(cw), _dp(ulo) if want_state else None)
        if rc != 0:
            raise RuntimeError("hifr_sgs_samples: " + lib().hifr_status_string(rc).decode())
        if want_state:
            return x, ce, cw, ulo.reshape((nl3, self.n_eles, 5), order="F")
        return x, ce, cw


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if shape is not None and a.shape != shape:
        raise ValueError("expected shape %s, got %s" % (shape, a.shape))
    return a


class MLP:
    """The learned eddy-viscosity network of include/hifr_sgs.h (device-resident:
    training, gradients and inference run in the library's HIP kernels)."""
    N_IN, N_PARAMS = 10, 9089

    def __init__(self, seed=0):
        self._h = C.c_void_p()
        self._check(lib().hifr_mlp_create(C.byref(self._h), int(seed)), "hifr_mlp_create")

    @staticmethod
    def _check(rc, what):
        if rc != 0:
            raise RuntimeError("%s: %s" % (what, lib().hifr_status_string(rc).decode()))

    def close(self):
        if self._h:
            lib().hifr_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def set_params(self, p):
        p = _f32(p, (self.N_PARAMS,))
        self._check(lib().hifr_mlp_set_params(self._h, _fp(p)), "hifr_mlp_set_params")

    def params(self):
        out = np.zeros(self.N_PARAMS, dtype=np.float32)
        self._check(lib().hifr_mlp_get_params(self._h, _fp(out)), "hifr_mlp_get_params")
        return out

    def set_normalization(self, mean, scale, lo, hi, out_scale):
        a = [_f32(v, (self.N_IN,)) for v in (mean, scale, lo, hi)]
        self._check(lib().hifr_mlp_set_normalization(self._h, *[_fp(v) for v in a], float(out_scale)),
                    "hifr_mlp_set_normalization")

    def loss_grad(self, x, t):
        x = _f32(x)
        t = _f32(t)
        n = t.size
        if x.shape != (n, self.N_IN):
            raise ValueError("x must be (n, 10)")
        loss = np.zeros(1, dtype=np.float32)
        g = np.zeros(self.N_PARAMS, dtype=np.float32)
        self._check(lib().hifr_mlp_loss_grad(self._h, n, _fp(x), _fp(t), _fp(loss), _fp(g)), "hifr_mlp_loss_grad")
        return float(loss[0]), g

    def set_data(self, x, t):
        x = _f32(x)
        t = _f32(t)
        if x.shape != (t.size, self.N_IN):
            raise ValueError("x must be (n, 10)")
        self._check(lib().hifr_mlp_set_data(self._h, t.size, _fp(x), _fp(t)), "hifr_mlp_set_data")

    def train(self, epochs, batch, lr):
        out = np.zeros(epochs, dtype=np.float32)
        self._check(lib().hifr_mlp_train(self._h, int(epochs), int(batch), float(lr), _fp(out)), "hifr_mlp_train")
        return out

    def predict(self, x):
        x = _f32(x)
        n = x.shape[0]
        if x.shape != (n, self.N_IN):
            raise ValueError("x must be (n, 10)")
        y = np.zeros(n, dtype=np.float32)
        self._check(lib().hifr_mlp_predict(self._h, n, _fp(x), _fp(y)), "hifr_mlp_predict")
        return y


def gauss_points(order):
    """The solver's 1-D Gauss-Legendre solution points of an order."""
    return np.polynomial.legendre.leggauss(order + 1)[0]


def modal_filter_1d(order_hi, order_lo):
    """(n_lo, n_hi) row-major: nodal values at the order_hi Gauss points -> the
    Legendre expansion truncated to degree order_lo, at the order_lo Gauss points."""
    xh, xl = gauss_points(order_hi), gauss_points(order_lo)
    modes = np.linalg.solve(np.polynomial.legendre.legvander(xh, order_hi), np.eye(order_hi + 1))
    return np.polynomial.legendre.legvander(xl, order_lo) @ modes[: order_lo + 1, :]


def lagrange_derivative_1d(x):
    """(n, n) row-major: D[a][b] = derivative of the Lagrange polynomial of point b at x[a]."""
    n = len(x)
    V = np.polynomial.legendre.legvander(x, n - 1)
    dV = np.stack([np.polynomial.legendre.legval(x, np.polynomial.legendre.legder(np.eye(n)[k]))
                   for k in range(n)], axis=1)
    return dV @ np.linalg.inv(V)
