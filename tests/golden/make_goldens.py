#!/usr/bin/env python3
"""Writes the golden residual/RK fixtures of tests/golden/*.npz from the reference
solver built by oracle/Makefile.ref (oracle/_ref/HiFiLES_cpu, serial, -D_NO_BLAS).

Runs only in the build container (it needs /root/reference for the cubature
tables and the reference binary); the fixtures it writes are plain data.

Protocol (state injection through the reference's restart reader):
  1. write a Gambit mesh, an input deck and Rest_000000000_p0000.dat holding a
     seeded state u0 printed with %.17g (exact doubles; the restart header holds
     the same 1-D points as the solver, so its interpolation opp_r is the identity);
  2. residual cases: one forward-Euler step (adv_type 0) with dt = DT_BIG; the
     restart written after the step holds u1 = u0 - dt*div/J to 15 digits, so
     R = (u0 - u1)/dt recovers div/J to ~1e-15 relative wherever |R| >> |u0|/dt;
  3. RK45 cases: n steps of adv_type 3 at the deck's dt; the final state is kept
     to the 15 digits the writer prints.

    python3 tests/golden/make_goldens.py        (from the repository root)
"""
import math
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import frmesh  # noqa: E402
from oracle import pyoracle as po  # noqa: E402

REF_HOME = "/root/reference"
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "HiFiLES_cpu")
DT_BIG = 1.0e10
TWO_PI = 6.2831853071795862


def deck_text(c, restart_iter=0):
    lines = [
        "equation 0", "viscous %d" % c["viscous"], "riemann_solve_type %d" % c.get("riemann", 0), "vis_riemann_solve_type 0",
        "ic_form %d" % c.get("ic_form", 1), "test_case 0", "order %d" % c["order"], "dt_type 0",
        "dt %.17g" % c["dt"], "n_steps %d" % c["n_steps"], "adv_type %d" % c["adv_type"],
        "tau %.17g" % c.get("tau", 0.0), "pen_fact %.17g" % c.get("pen_fact", 0.5),
        "restart_flag 1", "restart_iter %d" % restart_iter, "n_restart_files 1",
        "mesh_file mesh.%s" % c.get("mesh", "neu"),
        "dx_cyclic %.17g" % min(c["cyc"][0], 1e30), "dy_cyclic %.17g" % min(c["cyc"][1], 1e30),
        "dz_cyclic %.17g" % min(c["cyc"][2], 1e30),
        "p_res 2", "write_type 0", "inters_cub_order 3", "volume_cub_order %d" % c.get("vol_cub", 3),
        "monitor_res_freq 1", "plot_freq 1000000", "data_file_name golden",
        "restart_dump_freq %d" % c["n_steps"], "res_norm_type 1", "error_norm_type 1", "res_norm_field 0",
        "upts_type_quad 0", "vcjh_scheme_quad %d" % c["vcjh"], "eta_quad 0.0", "sparse_quad 0",
        "upts_type_hexa 0", "vcjh_scheme_hexa %d" % c["vcjh"], "eta_hexa 0.0", "sparse_hexa 0", "sparse_pri 0",
        "upts_type_tet 0", "fpts_type_tet 0", "vcjh_scheme_tet %d" % c["vcjh"], "c_tet 0.0", "sparse_tet 0",
        "gamma 1.4", "prandtl 0.72", "S_gas 120.", "T_gas 291.15", "R_gas 286.9", "mu_gas 1.827E-05",
        "fix_vis %d" % c.get("fix_vis", 1), "Mach_free_stream %.17g" % c.get("mach", 0.1), "nx_free_stream 1.",
        "ny_free_stream 0.", "nz_free_stream 0.", "Re_free_stream %.17g" % c.get("Re", 1600.0),
        "L_free_stream 1.", "T_free_stream 300", "T_wall %.17g" % c.get("T_wall", 300.0),
        "Mach_wall 0.", "nx_wall 0.", "ny_wall 0.", "nz_wall 0.",
        "rho_bound 1.", "u_bound 1.0", "v_bound 0.0", "w_bound 0.0", "p_bound 71.428571428571431",
    ]
    if c.get("integrals"):
        lines.append("integral_quantities %d %s" % (len(c["integrals"]), " ".join(c["integrals"])))
    if c.get("les"):
        lines += ["LES 1", "filter_type 0", "filter_ratio %.17g" % c["filter_ratio"],
                  "SGS_model %d" % c["sgs_model"], "wall_model 0", "wall_layer_thickness 0.0"]
    return "\n".join(lines) + "\n"


def oracle_deck(c):
    return po.default_deck(
        equation=0, viscous=c["viscous"], order=c["order"], adv_type=c["adv_type"], dt=c["dt"],
        tau=c.get("tau", 0.0), pen_fact=c.get("pen_fact", 0.5), ic_form=c.get("ic_form", 1),
        gamma=1.4, prandtl=0.72, Mach_free_stream=c.get("mach", 0.1), nx_free_stream=1.0,
        Re_free_stream=c.get("Re", 1600.0), L_free_stream=1.0, T_free_stream=300.0, T_wall=c.get("T_wall", 300.0),
        vcjh_scheme=c["vcjh"], eta=0.0, dx_cyclic=c["cyc"][0], dy_cyclic=c["cyc"][1],
        dz_cyclic=c["cyc"][2], LES=int(c.get("les", 0)), SGS_model=c.get("sgs_model", 0),
        filter_ratio=c.get("filter_ratio", 1.0), fix_vis=float(c.get("fix_vis", 1)),
        riemann_solve_type=c.get("riemann", 0))


def write_restart(path, ele_type, loc1d, order, u0, loc_upts=None):
    """u0: (Nu, N, nf) in the reference's point order; loc_upts (3, Nu) for tets."""
    nu, n, nf = u0.shape
    if ele_type == 2:  # src/eles_tets.cpp:749-800
        out = ["0", "TETS", "Order", "%d" % order, "Number of solution points per element", "%d" % nu,
               "Location of solution points in tetrahedral elements"]
        out += [" ".join("%.17g" % loc_upts[k, j] for k in range(3)) + " " for j in range(nu)]
        out += ["n_eles", "%d" % n, "ele2global_ele array", " ".join("%d" % i for i in range(n)) + " ", "data"]
    else:
        name = "HEXAS" if ele_type == 4 else "QUADS"
        what = "hexahedral" if ele_type == 4 else "quadrilateral"
        out = ["0", name, "Order", "%d" % order, "Number of solution points per %s element" % what,
               "%d" % nu, "Location of solution points in 1D", " ".join("%.17g" % v for v in loc1d) + " ",
               "n_eles", "%d" % n, "ele2global_ele array", " ".join("%d" % i for i in range(n)) + " ", "data"]
    for e in range(n):
        out.append("%d" % e)
        for j in range(nu):
            out.append(" ".join("%.17g" % u0[j, e, k] for k in range(nf)) + " ")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n\n")


def read_restart(path, nu, n, nf):
    with open(path) as fh:
        L = fh.read().split("\n")
    i = L.index("data") + 1
    u = np.zeros((nu, n, nf))
    for e in range(n):
        ele = int(L[i]); i += 1
        for j in range(nu):
            u[j, ele, :] = [float(v) for v in L[i].split()]
            i += 1
    return u


def write_restart_hifr(path, c, xv, conn_g, u0):
    """The initial restart file written by the library's own writer
    (hifr_solver_write_restart, host-only solver holding u0)."""
    sys.path.insert(0, os.path.join(ROOT, "hifiles-solver_amd"))
    import hifr
    c2v = frmesh.gambit_to_internal(c["ele_type"], conn_g)
    kw = dict(viscous=c["viscous"], order=c["order"], tau=c.get("tau", 0.0), pen_fact=c.get("pen_fact", 0.5),
              dt=c["dt"], adv_type=c["adv_type"], dx_cyclic=min(c["cyc"][0], 1e30),
              dy_cyclic=min(c["cyc"][1], 1e30), dz_cyclic=min(c["cyc"][2], 1e30), hifr_host_only=1,
              gamma=1.4, prandtl=0.72, Mach_free_stream=c.get("mach", 0.1), nx_free_stream=1.0,
              Re_free_stream=c.get("Re", 1600.0), L_free_stream=1.0, T_free_stream=300.0)
    kw[{4: "vcjh_scheme_hexa", 2: "vcjh_scheme_tet", 1: "vcjh_scheme_quad"}[c["ele_type"]]] = c["vcjh"]
    s = hifr.Solver.from_arrays(hifr.deck_text(**kw), c["ele_type"], xv, c2v)
    s.set_state(u0)
    s.write_restart(path)
    s.close()


def run_reference(c, xv, conn_g, u0, loc1d, loc_upts=None):
    wd = tempfile.mkdtemp(prefix="golden_")
    try:
        c2v = frmesh.gambit_to_internal(c["ele_type"], conn_g)
        walls = set(map(tuple, c.get("walls", [])))
        cyc = [(e, frmesh.internal_face_to_gambit(c["ele_type"], f))
               for e, f in frmesh.boundary_faces(c["ele_type"], c2v) if (e, f) not in walls]
        groups = [("Cyclic", cyc)]
        if walls:
            groups.append(("isotherm_fix", [(e, frmesh.internal_face_to_gambit(c["ele_type"], f))
                                            for e, f in sorted(walls)]))
        if c.get("mesh") == "msh":
            ggroups = [("CYCLIC", [(e, f) for e, f in frmesh.boundary_faces(c["ele_type"], c2v)
                                   if (e, f) not in walls])]
            if walls:
                ggroups.append(("isotherm_fix", sorted(walls)))
            frmesh.write_gmsh(os.path.join(wd, "mesh.msh"), xv, c2v, c["ele_type"], bc_groups=ggroups)
        else:
            frmesh.write_gambit(os.path.join(wd, "mesh.neu"), xv, conn_g, c["ele_type"], bc_groups=groups)
        with open(os.path.join(wd, "input"), "w") as fh:
            fh.write(deck_text(c))
        if c.get("restart_writer") == "hifr":
            write_restart_hifr(os.path.join(wd, "Rest_000000000_p0000.dat"), c, xv, conn_g, u0)
        else:
            write_restart(os.path.join(wd, "Rest_000000000_p0000.dat"), c["ele_type"], loc1d, c["order"], u0,
                          loc_upts)
        env = dict(os.environ, HIFILES_HOME=REF_HOME)
        r = subprocess.run([REF_BIN, "input"], cwd=wd, env=env, capture_output=True, text=True, timeout=3600)
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:] + r.stderr[-2000:])
        out = os.path.join(wd, "Rest_%09d_p0000.dat" % c["n_steps"])
        with open(out) as fh:
            c["rest_text"] = fh.read()
        if c.get("integrals"):
            with open(os.path.join(wd, "history.plt")) as fh:
                c["hist_text"] = fh.read()
        return read_restart(out, *u0.shape), r.stdout
    finally:
        shutil.rmtree(wd, ignore_errors=True)


INTEGRALS = ["kineticenergy", "vorticity", "pressuredilatation", "straincolonproduct", "devstraincolonproduct"]

CASES = [
    # name, element, sizes, order, scheme, physics, state
    dict(name="hex_p3_ns_tgv_2x2x2", ele_type=4, n=(2, 2, 2), order=3, vcjh=2, viscous=1, tau=1.0,
         pen_fact=0.5, state="tgv", noise=1e-3, seed=1234, perturb=0.0, kind="residual"),
    dict(name="hex_p3_ns_perturbed_3x3x3", ele_type=4, n=(3, 3, 3), order=3, vcjh=1, viscous=1, tau=0.5,
         pen_fact=0.5, state="tgv", noise=1e-2, seed=7, perturb=0.1, kind="residual"),
    dict(name="hex_p2_euler_3x2x2", ele_type=4, n=(3, 2, 2), order=2, vcjh=3, viscous=0,
         state="tgv", noise=1e-2, seed=11, perturb=0.08, kind="residual"),
    dict(name="hex_p4_ns_2x2x2", ele_type=4, n=(2, 2, 2), order=4, vcjh=2, viscous=1, tau=1.0,
         pen_fact=0.5, state="tgv", noise=1e-3, seed=5, perturb=0.0, kind="residual"),
    dict(name="quad_p3_euler_vortex_4x4", ele_type=1, n=(4, 4), order=3, vcjh=1, viscous=0,
         state="vortex", noise=0.0, seed=0, perturb=0.0, kind="residual"),
    dict(name="quad_p3_ns_perturbed_4x3", ele_type=1, n=(4, 3), order=3, vcjh=2, viscous=1, tau=1.0,
         pen_fact=0.5, state="vortex", noise=1e-3, seed=3, perturb=0.1, kind="residual"),
    dict(name="hex_p3_ns_rk45_2x2x2", ele_type=4, n=(2, 2, 2), order=3, vcjh=2, viscous=1, tau=1.0,
         pen_fact=0.5, state="tgv", noise=1e-3, seed=1234, perturb=0.0, kind="rk45", steps=3, dt=1e-3),
    # isothermal no-slip walls at y = 0 and y = 2 (channel), periodic in x and z
    dict(name="hex_p3_channel_walls_3x3x2", ele_type=4, n=(3, 3, 2), order=3, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="channel", noise=5e-2, seed=31, perturb=0.1, kind="residual", channel=True,
         mach=0.2, Re=2800.0, T_wall=310.0),
    dict(name="hex_p2_channel_walls_les_3x2x2", ele_type=4, n=(3, 2, 2), order=2, vcjh=2, viscous=1, tau=0.5,
         pen_fact=0.5, state="channel", noise=5e-2, seed=32, perturb=0.0, kind="residual", channel=True,
         mach=0.2, Re=2800.0, les=1, sgs_model=1, filter_ratio=2.0),
    dict(name="quad_p3_channel_walls_4x3", ele_type=1, n=(4, 3), order=3, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="channel", noise=5e-2, seed=33, perturb=0.1, kind="residual", channel=True,
         mach=0.2, Re=2800.0, T_wall=290.0),
    dict(name="hex_p3_channel_walls_rk45_2x2x2", ele_type=4, n=(2, 2, 2), order=3, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="channel", noise=5e-2, seed=34, perturb=0.0, kind="rk45", steps=3, dt=1e-3,
         channel=True, mach=0.2, Re=2800.0),
    # tetrahedra (hex boxes split into 6 tets), DG correction
    dict(name="tet_p3_ns_perturbed_2x2x2", ele_type=2, n=(2, 2, 2), order=3, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="tgv", noise=1e-2, seed=41, perturb=0.1, kind="residual"),
    dict(name="tet_p2_euler_2x2x2", ele_type=2, n=(2, 2, 2), order=2, vcjh=1, viscous=0,
         state="tgv", noise=1e-2, seed=42, perturb=0.0, kind="residual"),
    dict(name="tet_p1_ns_2x2x1", ele_type=2, n=(2, 2, 1), order=1, vcjh=1, viscous=1, tau=0.5,
         pen_fact=0.5, state="tgv", noise=1e-2, seed=43, perturb=0.0, kind="residual"),
    dict(name="tet_p3_channel_wale_walls_2x2x2", ele_type=2, n=(2, 2, 2), order=3, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="channel", noise=5e-2, seed=44, perturb=0.1, kind="residual", channel=True,
         mach=0.2, Re=2800.0, les=1, sgs_model=1, filter_ratio=2.0),
    dict(name="tet_p3_channel_rk45_2x1x1", ele_type=2, n=(2, 1, 1), order=3, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="channel", noise=5e-2, seed=45, perturb=0.0, kind="rk45", steps=3, dt=1e-3,
         channel=True, mach=0.2, Re=2800.0, les=1, sgs_model=1, filter_ratio=2.0),
    # LES: WALE / Smagorinsky eddy viscosity (reference CPU path)
    dict(name="hex_p3_les_wale_perturbed_3x3x3", ele_type=4, n=(3, 3, 3), order=3, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="tgv", noise=1e-2, seed=21, perturb=0.1, kind="residual", les=1, sgs_model=1,
         filter_ratio=2.0),
    dict(name="hex_p2_les_smagorinsky_3x2x2", ele_type=4, n=(3, 2, 2), order=2, vcjh=2, viscous=1, tau=0.5,
         pen_fact=0.5, state="tgv", noise=1e-2, seed=22, perturb=0.0, kind="residual", les=1, sgs_model=0,
         filter_ratio=1.5),
    dict(name="quad_p3_les_wale_perturbed_4x3", ele_type=1, n=(4, 3), order=3, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="vortex", noise=1e-3, seed=23, perturb=0.1, kind="residual", les=1, sgs_model=1,
         filter_ratio=2.0),
    dict(name="hex_p3_les_wale_rk45_2x2x2", ele_type=4, n=(2, 2, 2), order=3, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="tgv", noise=1e-2, seed=24, perturb=0.0, kind="rk45", steps=3, dt=1e-3, les=1,
         sgs_model=1, filter_ratio=2.0),
    # Sutherland's law (fix_vis 0, src/flux.cpp:326-329): the viscosity follows the
    # local temperature instead of the free-stream value
    dict(name="hex_p3_ns_sutherland_perturbed_2x2x2", ele_type=4, n=(2, 2, 2), order=3, vcjh=1, viscous=1,
         tau=1.0, pen_fact=0.5, state="tgv", noise=5e-2, seed=51, perturb=0.1, kind="residual", fix_vis=0),
    dict(name="quad_p3_ns_sutherland_perturbed_4x3", ele_type=1, n=(4, 3), order=3, vcjh=2, viscous=1, tau=1.0,
         pen_fact=0.5, state="vortex", noise=5e-2, seed=52, perturb=0.1, kind="residual", fix_vis=0),
    dict(name="tet_p2_ns_sutherland_2x2x1", ele_type=2, n=(2, 2, 1), order=2, vcjh=1, viscous=1, tau=0.5,
         pen_fact=0.5, state="tgv", noise=5e-2, seed=53, perturb=0.1, kind="residual", fix_vis=0),
    dict(name="hex_p2_channel_walls_sutherland_rk45_2x2x2", ele_type=4, n=(2, 2, 2), order=2, vcjh=1, viscous=1,
         tau=1.0, pen_fact=0.5, state="channel", noise=5e-2, seed=54, perturb=0.0, kind="rk45", steps=3, dt=1e-3,
         channel=True, mach=0.2, Re=2800.0, T_wall=320.0, fix_vis=0),
    # restart round trip: the initial state reaches the reference through the
    # library's restart writer; the reference's final restart file is kept as text
    dict(name="hex_p2_ns_restart_2x2x1", ele_type=4, n=(2, 2, 1), order=2, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="tgv", noise=1e-2, seed=81, perturb=0.0, kind="rk45", steps=2, dt=1e-3,
         restart_writer="hifr"),
    dict(name="tet_p2_ns_restart_2x1x1", ele_type=2, n=(2, 1, 1), order=2, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="tgv", noise=1e-2, seed=82, perturb=0.05, kind="rk45", steps=2, dt=1e-3,
         restart_writer="hifr"),
    dict(name="quad_p3_euler_restart_3x3", ele_type=1, n=(3, 3), order=3, vcjh=1, viscous=0,
         state="vortex", noise=1e-2, seed=83, perturb=0.05, kind="rk45", steps=2, dt=1e-3, restart_writer="hifr"),
    # Gmsh 2.2 meshes (mesh_file *.msh): the reference reads the same elements and
    # boundary groups through its Gmsh reader
    dict(name="hex_p3_ns_channel_gmsh_3x2x2", ele_type=4, n=(3, 2, 2), order=3, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="channel", noise=5e-2, seed=91, perturb=0.1, kind="residual", channel=True,
         mach=0.2, Re=2800.0, mesh="msh"),
    dict(name="tet_p2_ns_channel_gmsh_2x2x1", ele_type=2, n=(2, 2, 1), order=2, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="channel", noise=5e-2, seed=92, perturb=0.1, kind="residual", channel=True,
         mach=0.2, Re=2800.0, mesh="msh"),
    dict(name="quad_p3_ns_gmsh_perturbed_4x3", ele_type=1, n=(4, 3), order=3, vcjh=2, viscous=1, tau=1.0,
         pen_fact=0.5, state="vortex", noise=1e-2, seed=93, perturb=0.1, kind="residual", mesh="msh"),
    # integral diagnostics (CalcIntegralQuantities): history.plt of one forward-
    # Euler step with a negligible dt, so the state is u0 to the last bits
    dict(name="hex_p3_ns_integrals_3x3x3", ele_type=4, n=(3, 3, 3), order=3, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="tgv", noise=1e-2, seed=95, perturb=0.1, kind="integrals", vol_cub=4,
         integrals=INTEGRALS),
    dict(name="hex_p2_ns_integrals_2x2x2", ele_type=4, n=(2, 2, 2), order=2, vcjh=2, viscous=1, tau=1.0,
         pen_fact=0.5, state="tgv", noise=1e-2, seed=96, perturb=0.0, kind="integrals", vol_cub=6,
         integrals=INTEGRALS),
    dict(name="quad_p3_ns_integrals_4x3", ele_type=1, n=(4, 3), order=3, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="vortex", noise=1e-2, seed=97, perturb=0.1, kind="integrals", vol_cub=5,
         integrals=INTEGRALS),
    # Roe flux with entropy fix (riemann_solve_type 2, src/inters.cpp:390-520); the
    # reference evaluates it in 2-D only, so these are quads
    dict(name="quad_p3_euler_roe_perturbed_4x3", ele_type=1, n=(4, 3), order=3, vcjh=1, viscous=0,
         state="vortex", noise=1e-2, seed=61, perturb=0.1, kind="residual", riemann=2),
    dict(name="quad_p3_ns_roe_channel_walls_4x3", ele_type=1, n=(4, 3), order=3, vcjh=2, viscous=1, tau=1.0,
         pen_fact=0.5, state="channel", noise=5e-2, seed=62, perturb=0.1, kind="residual", channel=True,
         mach=0.2, Re=2800.0, T_wall=305.0, riemann=2),
    dict(name="quad_p2_ns_roe_rk45_3x3", ele_type=1, n=(3, 3), order=2, vcjh=1, viscous=1, tau=1.0,
         pen_fact=0.5, state="vortex", noise=1e-2, seed=63, perturb=0.05, kind="rk45", steps=3, dt=1e-3,
         riemann=2),
]


def read_cubature_table(path, nd):
    """{rule: (loc (d, n), weights (n))} of data/cubature_hexa.dat or cubature_quad.dat."""
    lines = open(path).read().split("\n")
    rules, i = {}, 0
    while i < len(lines):
        if lines[i].startswith("rule="):
            r = int(lines[i][5:])
            n = r ** nd
            i += 1
            cols = []
            for _ in range(nd + 1):
                cols.append([float(v) for v in lines[i + 1:i + 1 + n]])
                i += n + 1
            rules[r] = (np.array(cols[:nd]), np.array(cols[nd]))
        else:
            i += 1
    return rules


def cubature_rule(ele_type, rule):
    nd = 3 if ele_type == 4 else 2
    name = "cubature_hexa.dat" if ele_type == 4 else "cubature_quad.dat"
    return read_cubature_table(os.path.join(REF_HOME, "data", name), nd)[rule]


def write_cubature_fixture():
    """The reference's volume cubature tables as data: hex rules 2-8, quad rules 2-12."""
    rec = {}
    for ele_type, name, nd, rules in ((4, "cubature_hexa.dat", 3, range(2, 9)),
                                      (1, "cubature_quad.dat", 2, range(2, 13))):
        table = read_cubature_table(os.path.join(REF_HOME, "data", name), nd)
        for r in rules:
            rec["%s_%d_loc" % ("hex" if nd == 3 else "quad", r)] = table[r][0]
            rec["%s_%d_w" % ("hex" if nd == 3 else "quad", r)] = table[r][1]
    np.savez_compressed(os.path.join(HERE, "tables", "cubature_rules.npz"), **rec)


def build_case(c):
    if c["ele_type"] == 2:
        # hex box split into 6 tets per hex; channel walls or periodic
        nx, ny, nz = c["n"]
        if c.get("channel"):
            xv, hexes = frmesh.box_hex(nx, ny, nz, 2 * TWO_PI, 2.0, TWO_PI)
            c["cyc"] = (2 * TWO_PI, float("inf"), TWO_PI)
            h = 2.0 / ny
        else:
            xv, hexes = frmesh.box_hex(nx, ny, nz, TWO_PI, TWO_PI, TWO_PI)
            c["cyc"] = (TWO_PI, TWO_PI, TWO_PI)
            h = TWO_PI / max(c["n"])
        if c["perturb"] > 0:
            xv = frmesh.perturb_interior(xv, c["perturb"] * h, seed=c["seed"] + 100)
        conn = frmesh.hex_to_tets(xv, hexes)
        if c.get("channel"):
            c["walls"] = frmesh.wall_faces(2, conn, xv, 1, (0.0, 2.0))
        return xv, conn
    if c.get("channel"):
        # [0, 4pi] x [0, 2] (x [0, 2pi]), walls at y = 0, 2
        if c["ele_type"] == 4:
            nx, ny, nz = c["n"]
            xv, conn = frmesh.box_hex(nx, ny, nz, 2 * TWO_PI, 2.0, TWO_PI)
            c["cyc"] = (2 * TWO_PI, float("inf"), TWO_PI)
        else:
            nx, ny = c["n"]
            xv, conn = frmesh.box_quad(nx, ny, 2 * TWO_PI, 2.0)
            c["cyc"] = (2 * TWO_PI, float("inf"), float("inf"))
        h = 2.0 / c["n"][1]
        if c["perturb"] > 0:
            xv = frmesh.perturb_interior(xv, c["perturb"] * h, seed=c["seed"] + 100)
        c2v = frmesh.gambit_to_internal(c["ele_type"], conn)
        c["walls"] = frmesh.wall_faces(c["ele_type"], c2v, xv, 1, (0.0, 2.0))
        return xv, conn
    if c["ele_type"] == 4:
        nx, ny, nz = c["n"]
        xv, conn = frmesh.box_hex(nx, ny, nz, TWO_PI, TWO_PI, TWO_PI)
        c["cyc"] = (TWO_PI, TWO_PI, TWO_PI)
        h = TWO_PI / max(c["n"])
    else:
        nx, ny = c["n"]
        xv, conn = frmesh.box_quad(nx, ny, 20.0, 20.0, origin=(-10.0, -10.0))
        c["cyc"] = (20.0, 20.0, float("inf"))
        h = 20.0 / max(c["n"])
    if c["perturb"] > 0:
        xv = frmesh.perturb_interior(xv, c["perturb"] * h, seed=c["seed"] + 100)
    return xv, conn


def main():
    if not os.path.exists(REF_BIN):
        sys.exit("build the reference first: make -f oracle/Makefile.ref")
    only = sys.argv[1:]
    if not only or "cubature" in only:
        write_cubature_fixture()
    for c in CASES:
        if only and c["name"] not in only:
            continue
        xv, conn_g = build_case(c)
        c2v = frmesh.gambit_to_internal(c["ele_type"], conn_g)
        if c["kind"] == "residual":
            c.update(adv_type=0, dt=DT_BIG, n_steps=1)
        elif c["kind"] == "integrals":
            c.update(adv_type=0, dt=1e-16, n_steps=1)
        else:
            c.update(adv_type=3, n_steps=c["steps"])
        bfaces = [(e, f, 11) for e, f in c.get("walls", [])]
        o = po.Oracle(oracle_deck(c), c["ele_type"], xv, c2v, bfaces=bfaces or None)
        pos = o.get("pos_upts")
        if c["state"] == "tgv":
            u0 = frmesh.tgv_state(pos, p0=1.0 / (1.4 * 0.1 ** 2))
        elif c["state"] == "channel":
            u0 = frmesh.channel_state(pos, c["mach"])
        else:
            u0 = frmesh.shu_vortex_state(pos)
        if c["noise"] > 0:
            u0 = u0 * (1.0 + frmesh.seeded_perturbation(u0.shape, c["noise"], c["seed"]))
        # 17 digits reach the reference exactly; the library's writer prints 15
        fmt = "%.15g" if c.get("restart_writer") == "hifr" else "%.17g"
        u0 = np.array([float(fmt % v) for v in u0.ravel()]).reshape(u0.shape)
        loc1d = o.get("loc_upts")[0, : c["order"] + 1]
        u1, log = run_reference(c, xv, conn_g, u0, loc1d, o.get("loc_upts"))
        rec = dict(xv=xv, conn_gambit=conn_g, ele_type=c["ele_type"], order=c["order"],
                   vcjh=c["vcjh"], viscous=c["viscous"], tau=c.get("tau", 0.0),
                   pen_fact=c.get("pen_fact", 0.5), cyc=np.array(c["cyc"]), dt=c["dt"],
                   u0=u0, kind=c["kind"], les=int(c.get("les", 0)), sgs_model=int(c.get("sgs_model", 0)),
                   filter_ratio=float(c.get("filter_ratio", 1.0)),
                   bfaces=np.array(bfaces, dtype=np.int32).reshape(-1, 3), mach=c.get("mach", 0.1),
                   Re=c.get("Re", 1600.0), T_wall=c.get("T_wall", 300.0), fix_vis=int(c.get("fix_vis", 1)),
                   riemann=int(c.get("riemann", 0)))
        if c.get("restart_writer"):
            rec["rest_text"] = np.array(c["rest_text"])
        if c.get("mesh"):
            rec["mesh_format"] = np.array(c["mesh"])
        if c["kind"] == "integrals":
            # last line of history.plt: iteration, residual norms, forces, CL, CD,
            # the diagnostics, physical and computational time
            last = [v for v in c["hist_text"].strip().split("\n")][-1]
            vals = [float(v) for v in last.split(",")]
            nq = len(c["integrals"])
            rec["integral_names"] = np.array(c["integrals"])
            rec["integrals"] = np.array(vals[-2 - nq:-2])
            rec["vol_cub"] = c["vol_cub"]
            o.set_state(u0)
            o.residual()
            loc, w = cubature_rule(c["ele_type"], c["vol_cub"])
            got = o.integral_quantities(loc, w, list(range(nq)))
            rel = np.abs(got - rec["integrals"]) / np.abs(rec["integrals"])
            print("%-30s ref-vs-oracle integrals rel: %s" % (c["name"], " ".join("%.2e" % v for v in rel)))
        elif c["kind"] == "residual":
            R = (u0 - u1) / c["dt"]
            rec["R"] = R
            # oracle on the same mesh/state
            o.set_state(u0)
            o.residual()
            Ro = o.get("div_tconf_upts") / o.get("detjac_upts")[:, :, None]
            rel = np.abs(R - Ro).max(axis=(0, 1)) / np.abs(Ro).max(axis=(0, 1))
            # the oracle's own forward-Euler state printed as the reference prints it
            o.advance(0, 0)
            u1o = o.get("disu_upts")
            same = np.mean([float("%.15g" % a) == b for a, b in zip(u1o.ravel(), u1.ravel())])
            print("%-30s ref-vs-oracle max rel per field: %s; identical 15-digit states: %.4f" % (
                c["name"], " ".join("%.2e" % v for v in rel), same))
        else:
            rec["steps"] = c["steps"]
            rec["u_final"] = u1
            o.set_state(u0)
            o.zero_rk_register()
            o.run_steps(c["steps"])
            uo = o.get("disu_upts")
            rel = np.abs(u1 - uo).max(axis=(0, 1)) / np.abs(uo).max(axis=(0, 1))
            same = np.mean([float("%.15g" % a) == b for a, b in zip(uo.ravel(), u1.ravel())])
            print("%-30s ref-vs-oracle RK45 state max rel: %s; identical 15-digit states: %.4f" % (
                c["name"], " ".join("%.2e" % v for v in rel), same))
        np.savez_compressed(os.path.join(HERE, c["name"] + ".npz"), **rec)


if __name__ == "__main__":
    main()
