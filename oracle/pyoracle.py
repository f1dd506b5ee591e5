"""ctypes binding of oracle/liboracle.so (the CPU checker).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
bench's CPU-baseline leg, never by the product.  Arrays follow the reference's
column-major layouts; numpy arrays returned here are Fortran-ordered views with
the reference's index order, e.g. disu_upts[j, i, k] = disu_upts(j,i,k).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Deck(C.Structure):
    """or_deck in fr_oracle.h (raw input-deck values, src/input.cpp:89-331)."""
    _fields_ = [
        ("equation", C.c_int), ("viscous", C.c_int), ("riemann_solve_type", C.c_int),
        ("vis_riemann_solve_type", C.c_int), ("ic_form", C.c_int), ("order", C.c_int),
        ("adv_type", C.c_int), ("dt", C.c_double), ("tau", C.c_double), ("pen_fact", C.c_double),
        ("gamma", C.c_double), ("prandtl", C.c_double), ("S_gas", C.c_double), ("T_gas", C.c_double),
        ("R_gas", C.c_double), ("mu_gas", C.c_double), ("fix_vis", C.c_double),
        ("Mach_free_stream", C.c_double), ("nx_free_stream", C.c_double),
        ("ny_free_stream", C.c_double), ("nz_free_stream", C.c_double),
        ("Re_free_stream", C.c_double), ("L_free_stream", C.c_double),
        ("T_free_stream", C.c_double), ("T_wall", C.c_double), ("Mach_wall", C.c_double),
        ("nx_wall", C.c_double), ("ny_wall", C.c_double), ("nz_wall", C.c_double),
        ("upts_type", C.c_int), ("vcjh_scheme", C.c_int), ("eta", C.c_double),
        ("LES", C.c_int), ("SGS_model", C.c_int), ("filter_ratio", C.c_double),
        ("const_src", C.c_double), ("dx_cyclic", C.c_double), ("dy_cyclic", C.c_double),
        ("dz_cyclic", C.c_double),
    ]


class Phys(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "gamma", "prandtl", "prandtl_t", "mu_inf", "rt_inf", "c_sth", "R_ref", "fix_vis",
        "tau", "pen_fact", "T_wall", "const_src", "dt")] + [("v_wall", C.c_double * 3)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.or_deck_defaults.argtypes = [C.POINTER(Deck)]
        L.or_setup_params.argtypes = [C.POINTER(Deck), C.POINTER(Phys)]
        L.or_create.argtypes = [C.POINTER(Deck), C.c_int, C.c_int, dp, C.c_int, ip]
        L.or_create.restype = vp
        L.or_create_bc.argtypes = [C.POINTER(Deck), C.c_int, C.c_int, dp, C.c_int, ip, C.c_int, ip, ip, ip]
        L.or_create_bc.restype = vp
        L.or_destroy.argtypes = [vp]
        for f in ("or_n_upts", "or_n_fpts", "or_n_fields", "or_n_dims", "or_n_eles", "or_n_inters"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = C.c_int
        L.or_get_phys.argtypes = [vp, C.POINTER(Phys)]
        L.or_get.argtypes = [vp, C.c_int, dp]
        L.or_get.restype = C.c_long
        L.or_set_state.argtypes = [vp, dp]
        L.or_set_sgs_nu.argtypes = [vp, dp]
        L.or_set_metrics.argtypes = [vp, dp, dp, dp, dp]
        L.or_zero_rk_register.argtypes = [vp]
        L.or_set_ics_tgv.argtypes = [vp, C.c_double]
        L.or_get_face_pairs.argtypes = [vp, ip, ip]
        L.or_get_face_pairs.restype = C.c_long
        L.or_residual.argtypes = [vp]
        L.or_advance.argtypes = [vp, C.c_int, C.c_int]
        L.or_run_steps.argtypes = [vp, C.c_int]
        L.or_norm_residual.argtypes = [vp, C.c_int, dp]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_common_flux.argtypes = [C.c_int, C.c_int, C.c_double, dp, dp, dp, dp]
        L.or_integral_quantities.argtypes = [vp, C.c_int, dp, dp, C.c_int, ip, dp]
        _LIB = L
    return _LIB


def default_deck(**kw):
    d = Deck()
    lib().or_deck_defaults(C.byref(d))
    for k, v in kw.items():
        if not hasattr(d, k):
            raise KeyError(k)
        setattr(d, k, v)
    return d


def common_flux(riemann, ul, ur, n, gamma=1.4):
    """Common normal flux of one point (riemann 0 Rusanov, 2 Roe), static mesh."""
    ul, ur, n = (np.ascontiguousarray(v, dtype=np.float64) for v in (ul, ur, n))
    fn = np.zeros(len(ul))
    lib().or_common_flux(len(n), riemann, gamma, _dp(ul), _dp(ur), _dp(n), _dp(fn))
    return fn


def setup_params(deck):
    p = Phys()
    lib().or_setup_params(C.byref(deck), C.byref(p))
    return p


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One element type on one mesh.  ele_type: 1 quad, 4 hex (reference numbering)."""

    WHAT = {"disu_upts": 0, "rk_register": 1, "div_tconf_upts": 2, "disu_fpts": 3,
            "grad_disu_upts": 4, "grad_disu_fpts": 5, "detjac_upts": 6, "JGinv_upts": 7,
            "tdA_fpts": 8, "norm_fpts": 9, "pos_upts": 10, "norm_tconf_fpts": 11,
            "opp_0": 12, "opp_3": 13, "opp_4": 14, "loc_upts": 15, "tloc_fpts": 16,
            "pos_fpts": 17, "sgsf_upts": 18, "sgsf_fpts": 19}

    def __init__(self, deck, ele_type, xv, c2v, threads=None, bfaces=None):
        """bfaces: optional (n, 3) int array of (element, local face, bctype) boundary faces."""
        L = lib()
        if threads is not None:
            L.or_set_threads(int(threads))
        self.deck = deck
        xv = np.ascontiguousarray(xv, dtype=np.float64)
        c2v = np.ascontiguousarray(c2v, dtype=np.int32)
        bf = np.zeros((0, 3), dtype=np.int32) if bfaces is None else np.asarray(bfaces, dtype=np.int32).reshape(-1, 3)
        be, bfc, bt = (np.ascontiguousarray(bf[:, i]) for i in range(3))
        ip = C.POINTER(C.c_int)
        self._h = L.or_create_bc(C.byref(deck), int(ele_type), xv.shape[0], _dp(xv), c2v.shape[0],
                                 c2v.ctypes.data_as(ip), bf.shape[0], be.ctypes.data_as(ip),
                                 bfc.ctypes.data_as(ip), bt.ctypes.data_as(ip))
        if not self._h:
            raise RuntimeError("or_create failed")
        self.nu = L.or_n_upts(self._h)
        self.nfp = L.or_n_fpts(self._h)
        self.nf = L.or_n_fields(self._h)
        self.nd = L.or_n_dims(self._h)
        self.n_eles = L.or_n_eles(self._h)
        self.n_inters = L.or_n_inters(self._h)

    def close(self):
        if self._h:
            lib().or_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape_of(self, name):
        nu, nfp, N, nf, nd = self.nu, self.nfp, self.n_eles, self.nf, self.nd
        return {"disu_upts": (nu, N, nf), "rk_register": (nu, N, nf), "div_tconf_upts": (nu, N, nf),
                "disu_fpts": (nfp, N, nf), "grad_disu_upts": (nu, N, nf, nd),
                "grad_disu_fpts": (nfp, N, nf, nd), "detjac_upts": (nu, N),
                "JGinv_upts": (nd, nd, nu, N), "tdA_fpts": (nfp, N), "norm_fpts": (nfp, N, nd),
                "pos_upts": (nu, N, nd), "norm_tconf_fpts": (nfp, N, nf), "opp_0": (nfp, nu),
                "opp_3": (nu, nfp), "opp_4": (nu, nu, nd), "loc_upts": (nd, nu),
                "tloc_fpts": (nd, nfp), "pos_fpts": (nfp, N, nd), "sgsf_upts": (nu, N, nf, nd),
                "sgsf_fpts": (nfp, N, nf, nd)}[name]

    def get(self, name):
        shp = self.shape_of(name)
        out = np.zeros(int(np.prod(shp)), dtype=np.float64)
        n = lib().or_get(self._h, self.WHAT[name], _dp(out))
        if n != out.size:
            raise RuntimeError("or_get(%s) returned %d" % (name, n))
        return out.reshape(shp, order="F")

    def set_state(self, u):
        u = np.asarray(u, dtype=np.float64)
        flat = np.ascontiguousarray(u.reshape(-1, order="F"))
        if flat.size != self.nu * self.n_eles * self.nf:
            raise ValueError("state size")
        lib().or_set_state(self._h, _dp(flat))

    def set_ics_tgv(self, rho=1.0):
        lib().or_set_ics_tgv(self._h, float(rho))

    def zero_rk_register(self):
        lib().or_zero_rk_register(self._h)

    def face_pairs(self):
        n = lib().or_get_face_pairs(self._h, None, None)
        left = np.zeros(n, dtype=np.int32)
        right = np.zeros(n, dtype=np.int32)
        lib().or_get_face_pairs(self._h, left.ctypes.data_as(C.POINTER(C.c_int)),
                                right.ctypes.data_as(C.POINTER(C.c_int)))
        return left, right

    def residual(self):
        lib().or_residual(self._h)
        return self.get("div_tconf_upts")

    def advance(self, stage, adv_type):
        lib().or_advance(self._h, int(stage), int(adv_type))

    def set_metrics(self, detjac_upts, JGinv_upts, tdA_fpts, norm_fpts):
        """Replaces the oracle's own metrics (reference layouts, Fortran order)."""
        arrs = []
        for name, a in (("detjac_upts", detjac_upts), ("JGinv_upts", JGinv_upts), ("tdA_fpts", tdA_fpts),
                        ("norm_fpts", norm_fpts)):
            a = np.asarray(a, dtype=np.float64)
            if a.shape != self.shape_of(name):
                raise ValueError("%s: shape %s, expected %s" % (name, a.shape, self.shape_of(name)))
            arrs.append(np.ascontiguousarray(a.reshape(-1, order="F")))
        lib().or_set_metrics(self._h, *[_dp(a) for a in arrs])

    def set_sgs_nu(self, nu_t):
        """SGS_model 10: kinematic eddy viscosity at the solution points, (Nu, N)."""
        flat = np.ascontiguousarray(np.asarray(nu_t, dtype=np.float64).reshape(-1, order="F"))
        if flat.size != self.nu * self.n_eles:
            raise ValueError("nu_t size")
        lib().or_set_sgs_nu(self._h, _dp(flat))

    def integral_quantities(self, loc, weights, kinds):
        """Integral diagnostics of the current state with the corrected gradient of
        the last residual(); loc (d, ncub), weights (ncub)."""
        loc = np.ascontiguousarray(np.asarray(loc, dtype=np.float64).reshape(-1, order="F"))
        w = np.ascontiguousarray(weights, dtype=np.float64)
        k = np.ascontiguousarray(kinds, dtype=np.int32)
        out = np.zeros(len(k))
        lib().or_integral_quantities(self._h, len(w), _dp(loc), _dp(w), len(k),
                                     k.ctypes.data_as(C.POINTER(C.c_int)), _dp(out))
        return out

    def run_steps(self, n):
        lib().or_run_steps(self._h, int(n))

    def norms(self, norm_type=1):
        out = np.zeros(5)
        lib().or_norm_residual(self._h, int(norm_type), _dp(out))
        return out[: self.nf]

    def phys(self):
        p = Phys()
        lib().or_get_phys(self._h, C.byref(p))
        return p
