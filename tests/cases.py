"""Test cases shared by the CPU and GPU tests: the golden fixtures of
tests/golden/*.npz and the decks that reproduce them for the oracle and for the
GPU solver."""
import glob
import os

import numpy as np

import frmesh

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TWO_PI = 6.2831853071795862

# physical parameters of tests/golden/make_goldens.py decks
PHYS = dict(gamma=1.4, prandtl=0.72, Mach_free_stream=0.1, nx_free_stream=1.0, Re_free_stream=1600.0,
            L_free_stream=1.0, T_free_stream=300.0, T_wall=300.0)


def golden_names(kind=None):
    names = sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz")))
    if kind is None:
        return names
    return [n for n in names if str(load_golden(n)["kind"]) == kind]


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def mesh_of(g):
    ele_type = int(g["ele_type"])
    return ele_type, g["xv"], frmesh.gambit_to_internal(ele_type, g["conn_gambit"])


def bfaces_of(g):
    """(n, 3) boundary faces (element, local face, bctype) of a fixture, or None."""
    bf = g.get("bfaces")
    return bf if bf is not None and len(bf) else None


def phys_of(g):
    """Physical deck parameters of a fixture (Mach, Re, T_wall may differ)."""
    ph = dict(PHYS)
    for key, name in (("mach", "Mach_free_stream"), ("Re", "Re_free_stream"), ("T_wall", "T_wall")):
        if key in g:
            ph[name] = float(g[key])
    return ph


def solver_of(g, **over):
    import hifr
    ele_type, xv, c2v = mesh_of(g)
    return hifr.Solver.from_arrays(product_deck(g, **over), ele_type, xv, c2v, bfaces=bfaces_of(g))


def oracle_of(g, threads=None, **over):
    from oracle import pyoracle as po
    ele_type, xv, c2v = mesh_of(g)
    return po.Oracle(oracle_deck(g, **over), ele_type, xv, c2v, threads=threads, bfaces=bfaces_of(g))


def product_deck(g, host_only=False, **over):
    import hifr
    ele_type = int(g["ele_type"])
    cyc = [float(v) for v in g["cyc"]]
    kw = dict(viscous=int(g["viscous"]), order=int(g["order"]), tau=float(g["tau"]),
              pen_fact=float(g["pen_fact"]), dt=float(g["dt"]),
              adv_type=3 if str(g["kind"]) == "rk45" else 0,
              dx_cyclic=cyc[0], dy_cyclic=cyc[1], dz_cyclic=cyc[2], fix_vis=int(g.get("fix_vis", 1)),
              riemann_solve_type=int(g.get("riemann", 0)))
    kw.update(phys_of(g))
    if ele_type == 4:
        kw["vcjh_scheme_hexa"] = int(g["vcjh"])
    elif ele_type == 2:
        kw["vcjh_scheme_tet"] = int(g["vcjh"])
    else:
        kw["vcjh_scheme_quad"] = int(g["vcjh"])
    if int(g.get("les", 0)):
        kw.update(LES=1, SGS_model=int(g["sgs_model"]), filter_ratio=float(g["filter_ratio"]), filter_type=0,
                  wall_model=0)
    if host_only:
        kw["hifr_host_only"] = 1
    kw.update(over)
    return hifr.deck_text(**kw)


def oracle_deck(g, **over):
    from oracle import pyoracle as po
    cyc = [float(v) for v in g["cyc"]]
    kw = dict(equation=0, viscous=int(g["viscous"]), order=int(g["order"]),
              adv_type=3 if str(g["kind"]) == "rk45" else 0, dt=float(g["dt"]), tau=float(g["tau"]),
              pen_fact=float(g["pen_fact"]), vcjh_scheme=int(g["vcjh"]), eta=0.0,
              dx_cyclic=cyc[0], dy_cyclic=cyc[1], dz_cyclic=cyc[2], LES=int(g.get("les", 0)),
              SGS_model=int(g.get("sgs_model", 0)), filter_ratio=float(g.get("filter_ratio", 1.0)),
              fix_vis=float(g.get("fix_vis", 1)), riemann_solve_type=int(g.get("riemann", 0)))
    kw.update(phys_of(g))
    kw.update(over)
    return po.default_deck(**kw)


def field_rel_err(a, b):
    """max |a-b| / max |b| per conserved variable (last axis of a (.., nf) array)."""
    a = np.asarray(a)
    b = np.asarray(b)
    nf = b.shape[-1] if b.ndim == 3 else b.shape[2]
    out = []
    for k in range(nf):
        bk = b[..., k] if b.ndim == 3 else b[:, :, k]
        ak = a[..., k] if a.ndim == 3 else a[:, :, k]
        out.append(float(np.abs(ak - bk).max() / max(np.abs(bk).max(), 1e-300)))
    return np.array(out)


def channel_tet_case(n=(24, 12, 12), order=3, perturb=0.1, mesh_seed=2024, dt=1e-3, steps=200, sgs_model=1,
                     filter_ratio=2.0):
    """The unstructured tetrahedral channel: an n[0] x n[1] x n[2] hex box over
    [0,4pi] x [0,2] x [0,2pi], each hex split into 6 tets, interior vertices moved
    by up to `perturb` of the y spacing, isothermal no-slip walls at y = 0 and 2,
    periodic in x and z; Mach 0.2, Re_bulk 2800, WALE (sgs_model 1): a golden-like
    record for solver_of / oracle_of (channel_tet_state gives the initial field)."""
    xv, hexes = frmesh.box_hex(n[0], n[1], n[2], 2 * TWO_PI, 2.0, TWO_PI)
    if perturb > 0:
        xv = frmesh.perturb_interior(xv, perturb * 2.0 / n[1], seed=mesh_seed)
    conn = frmesh.hex_to_tets(xv, hexes)
    walls = frmesh.wall_faces(2, conn, xv, 1, (0.0, 2.0))
    g = dict(ele_type=2, xv=xv, conn_gambit=conn, order=order, vcjh=1, viscous=1, tau=1.0, pen_fact=0.5,
             cyc=np.array([2 * TWO_PI, float("inf"), TWO_PI]), dt=dt, kind="rk45", steps=steps, les=1,
             sgs_model=sgs_model, filter_ratio=filter_ratio,
             bfaces=np.array([(e, f, 11) for e, f in walls], dtype=np.int32).reshape(-1, 3), mach=0.2,
             Re=2800.0, T_wall=300.0)
    return g


def channel_tet_state(g, pos, seed=7, noise=5e-2):
    """Parabolic channel profile (rho = 1, uniform pressure) plus seeded random
    velocity noise of `noise` times the bulk velocity, damped as y (2 - y) so that
    the walls stay at rest; the energy carries the perturbed kinetic energy."""
    u0 = frmesh.channel_state(pos, float(g["mach"]))
    y = pos[..., 1]
    dv = frmesh.seeded_perturbation(pos.shape, noise, seed) * (y * (2.0 - y))[..., None]
    ke0 = 0.5 * (u0[..., 1] ** 2 + u0[..., 2] ** 2 + u0[..., 3] ** 2) / u0[..., 0]
    u0[..., 1:4] += u0[..., :1] * dv
    ke1 = 0.5 * (u0[..., 1] ** 2 + u0[..., 2] ** 2 + u0[..., 3] ** 2) / u0[..., 0]
    u0[..., 4] += ke1 - ke0
    return u0
