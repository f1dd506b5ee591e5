#!/usr/bin/env python3
"""Throughput of the FR residual + RK45 update on one MI355X, FP64 and mixed precision.

Default workload (the headline configuration): 3-D compressible Navier-Stokes,
Taylor-Green initial field (Mach 0.1, Re 1600, Pr 0.71) plus a seeded 1e-3
perturbation, periodic cube [0,2pi]^3 with 32^3 hexes at P=3 (64 solution points
per element, 10,485,760 DOF), Rusanov + LDG, RK45, dt = 1e-3, FP64.  A step is one
RK45 step (5 stages, i.e. 5 residual evaluations + updates); the metric counts
DOF x stages per second.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
                  [--workload tgv32 | tgv48-p4 | tet-channel | vortex-quads]

Other workloads (same JSON line, for the record of each configuration):
  tgv48-p4      the same flow on 48^3 hexes at P=4 (69.1 M DOF), dt = 5e-4;
  tet-channel   24 x 12 x 12 hexes split into 6 tets each (20,736 tets, P=3,
                2.07 M DOF), 10% vertex perturbation, isothermal no-slip walls,
                WALE, Mach 0.2, Re_bulk 2800, dt = 1e-3;
  vortex-quads  2-D Euler isentropic vortex (Mach 0.5) on 20^2 periodic quads
                on [-10,10]^2, P=3 (25,600 DOF), Rusanov, dt = 5e-3: a mesh
                small enough that launches, not bytes, set the rate, so the line
                adds `graph_replay`, the rate of hifr_run's graph replays.

Prints one JSON line.  `value` is the FP64 path; `mixed` holds the same
measurement of the mixed-precision path (FP32 arithmetic and traces, FP64 state
and RK update) with its error against the FP64 state after the same steps.
`roofline` is for the dominant kernel (flux, divergence
and RK update, one launch per stage): algorithmic bytes per launch / its average
duration from HIP events recorded around the launches of the first stage of every
timed step (events around every launch would lengthen the loop by about 9 %).
`stage.grad_kernel_roofline` is the same figure for the gradient-trace kernel.
`per_point_fp64` repeats the FP64 measurement with per-point metrics (the storage a
curved mesh needs; the default keeps one metric set per affine element).
`cpu_baseline` is the oracle (the C restatement of the reference's dense-operator
CPU residual) on the same mesh and state on this host: all physical cores (median
of 5 RK45 steps) and one core (median of 5 RK stages), with the CPU model and the
oracle's calibration against the reference binary.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hifiles-solver_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import hifr  # noqa: E402

METRIC = "GDOF-RK-stages/s, 3D NS hex P=3, FP64 vs mixed; % of HBM roofline"
HBM_PEAK_GBS = 8000.0
TWO_PI = 6.2831853071795862


def deck(order, n_steps, metrics=0, precision=0, dt=1e-3):
    return hifr.deck_text(hifr_metrics=metrics, hifr_precision=precision,
                          viscous=1, order=order, vcjh_scheme_hexa=1, dt=dt, adv_type=3, tau=1.0, pen_fact=0.5,
                          gamma=1.4, prandtl=0.71, Mach_free_stream=0.1, nx_free_stream=1.0,
                          Re_free_stream=1600.0, L_free_stream=1.0, T_free_stream=300.0, fix_vis=1,
                          dx_cyclic=TWO_PI, dy_cyclic=TWO_PI, dz_cyclic=TWO_PI, n_steps=n_steps)


def tgv_state(pos, seed=1234):
    import frmesh
    u = frmesh.tgv_state(pos, p0=1.0 / (1.4 * 0.1 ** 2))
    return u * (1.0 + frmesh.seeded_perturbation(u.shape, 1e-3, seed))


WORKLOADS = {
    "tgv32": dict(kind="tgv", n=32, order=3, dt=1e-3),
    "tgv48-p4": dict(kind="tgv", n=48, order=4, dt=5e-4),
    "tet-channel": dict(kind="tet", order=3, dt=1e-3),
    "vortex-quads": dict(kind="vortex", n=20, order=3, dt=5e-3),
}
VORTEX_L = 20.0


def describe(w, n, order):
    if w["kind"] == "vortex":
        return ("2-D Euler isentropic vortex (beta 5, Mach 0.5 advection), %d^2 periodic quads on [-10,10]^2, P=%d, "
                "Rusanov, RK45 dt=%g, FP64" % (n, order, w["dt"]))
    if w["kind"] == "tet":
        return ("3-D Navier-Stokes + WALE plane channel (Mach 0.2, Re_bulk 2800), 24x12x12 hexes split into "
                "20,736 tets with 10%% vertex perturbation, P=%d, isothermal no-slip walls, Rusanov+LDG, "
                "RK45 dt=%g, FP64" % (order, w["dt"]))
    return ("3-D Navier-Stokes Taylor-Green (Mach 0.1, Re 1600), %d^3 periodic hexes, P=%d, Rusanov+LDG, "
            "RK45 dt=%g, FP64" % (n, order, w["dt"]))


def make_solver(w, n, order, steps, per_point, precision):
    """Solver with the workload's initial state uploaded; the oracle's inputs."""
    prec = int(precision == "mixed")
    if w["kind"] == "vortex":
        import frmesh
        xv, conn = frmesh.box_quad(n, n, VORTEX_L, VORTEX_L, origin=(-10.0, -10.0))
        c2v = frmesh.gambit_to_internal(1, conn)
        d = hifr.deck_text(hifr_metrics=int(per_point), hifr_precision=prec, viscous=0, order=order,
                           vcjh_scheme_quad=1, dt=w["dt"], adv_type=3, dx_cyclic=VORTEX_L, dy_cyclic=VORTEX_L,
                           n_steps=steps)
        s = hifr.Solver.from_arrays(d, 1, xv, c2v)
        s.set_state(frmesh.shu_vortex_state(s.get("pos_upts")))
        return s, (xv, c2v)
    if w["kind"] == "tet":
        import cases
        g = cases.channel_tet_case(order=order, dt=w["dt"])
        s = cases.solver_of(g, hifr_metrics=int(per_point), hifr_precision=prec)
        s.set_state(cases.channel_tet_state(g, s.get("pos_upts")))
        return s, g
    s = hifr.Solver.box(deck(order, steps, int(per_point), prec, w["dt"]), hifr.HEX, (n, n, n), (TWO_PI,) * 3,
                        (0.0,) * 3)
    s.set_state(tgv_state(s.get("pos_upts")))
    return s, None


def algorithmic_bytes(nu, nfp, nfaces, nd=3, per_element=False, mixed=False, viscous=True):
    """HBM bytes per element that the two kernels of a Navier-Stokes stage must
    move, counting every array once:
      gradient-trace kernel: read u, neighbour u_f, JGinv, detJ, face normals,
                             neighbour index; write the viscous normal flux
                             traces F_v.n (nf per flux point; with LES the SGS
                             normal flux is folded into the same traces);
      flux/RK kernel:        read u, RK register, neighbour u_f, own and neighbour
                             F_v.n, JGinv, detJ, face normal + tdA, neighbour
                             index; write u, RK register, next u_f.
    u, RK register, detJ, normals and tdA are FP64; traces and JGinv are FP64, or
    FP32 on the mixed path; the neighbour index is int32.  Metrics are per
    solution point / flux point, or with per_element (every element affine) one
    JGinv + detJ per element and one normal + tdA per face.  Euler (not viscous):
    no gradient-trace kernel, and the flux/RK kernel reads no F_v.n traces."""
    nf = nd + 2
    t = 4 if mixed else 8
    nv = 1 if per_element else nu
    vol = t * nd * nd * nv + 8 * nv
    fpts = nfaces if per_element else nfp
    grad = 8 * nu * nf + t * nfp * nf + vol + 8 * fpts * nd + t * nfp * nf + 4 * nfp
    rk = 8 * 4 * nu * nf + t * nfp * nf + t * 2 * nfp * nf + vol + 8 * fpts * (nd + 1) + t * nfp * nf + 4 * nfp
    if not viscous:
        return 0, rk - t * 2 * nfp * nf
    return grad, rk


# Where the traffic constants below come from: they are not measured by this run.
TRAFFIC_SOURCE = {
    "tgv32": "constant from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE in profiles/fr_tensor_kernel_tgv32_p3.md",
    "tgv48-p4": "constant from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE in profiles/fr_tensor_kernel_tgv48_p4.md",
    "tet-channel": "constant from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE in profiles/fr_tet_mfma_kernel_tet_channel_p3.md",
}

# HBM traffic per element of each stage kernel from rocprofv3 --pmc FETCH_SIZE (x2.0)
# and WRITE_SIZE (x1.0), KiB per dispatch, the factors read off a 1 GiB copy kernel;
# keyed by (workload, per-element metrics, precision).  Sources: the summaries
# under profiles/ named for each kernel and workload.
MEASURED_KIB = {
    ("tgv32", False, "fp64"): {"n_eles": 32768, "grad": (299091.6, 122882.3), "rk": (445657.0, 297643.6)},
    ("tgv32", True, "fp64"): {"n_eles": 32768, "grad": (153924.2, 122881.0), "rk": (303791.0, 305837.2)},
    ("tgv32", True, "mixed"): {"n_eles": 32768, "grad": (105990.1, 61440.3), "rk": (245266.1, 236204.0)},
    ("tgv48-p4", False, "fp64"): {"n_eles": 110592, "grad": (2011828.6, 653347.8), "rk": (2915802.9, 1847838.9)},
    ("tgv48-p4", True, "fp64"): {"n_eles": 110592, "grad": (1052445.8, 651318.4), "rk": (1916025.3, 1848559.3)},
    ("tgv48-p4", True, "mixed"): {"n_eles": 110592, "grad": (739545.1, 328894.9), "rk": (1708124.0, 1515354.9)},
    ("tet-channel", False, "fp64"): {"n_eles": 20736, "grad": (92341.8, 32400.6), "rk": (137424.9, 66960.5)},
    ("tet-channel", True, "fp64"): {"n_eles": 20736, "grad": (54497.2, 32400.1), "rk": (105189.8, 66960.6)},
    ("tet-channel", True, "mixed"): {"n_eles": 20736, "grad": (45626.8, 16200.1), "rk": (92839.0, 50760.1)},
}


# hifr_run steps meshes of at most this many solution points by graph replay
GRAPH_MAX_POINTS = 16384

COPY_DOUBLES = 1 << 27  # 1 GiB read and 1 GiB written per copy


def copy_ms(n_doubles, reps=10):
    """Average time of a plain device copy of n_doubles FP64 values
    (hifr_calibrate_copy)."""
    import ctypes as C
    L = hifr.lib()
    L.hifr_calibrate_copy.argtypes = [C.c_long, C.c_int, C.POINTER(C.c_double)]
    ms = C.c_double()
    rc = L.hifr_calibrate_copy(n_doubles, reps, C.byref(ms))
    if rc != 0:
        raise SystemExit("copy calibration failed (%d)" % rc)
    return ms.value


def measured_traffic(workload, per_element, precision):
    m = MEASURED_KIB.get((workload, per_element, precision))
    if not m:
        return None
    return {k: (m[k][0] * 2.0 + m[k][1]) * 1024 / m["n_eles"] for k in ("grad", "rk")}


def survey_model_bytes(nu, nfp, nd=3, les=False):
    """Bytes per element of one Navier-Stokes stage in the three-phase model
    (u -> u_f; u + u_f -> grad u_f; u + u_f + grad u_f -> flux, div, RK) that
    counts every operand once per phase; this schedule moves fewer bytes."""
    nf = nd + 2
    b = 8 * (3 * nu * nf + nu * nf + 2 * nu * nf + 3 * nfp * nf + 2 * nfp * nf * nd + 2 * nu * (1 + nd * nd)
             + 2 * nfp * (nd + 1))
    return b + (8 * 2 * nfp * nf * nd if les else 0)


def cpu_model():
    """CPU model name and the physical cores this process may run on."""
    model, cores = "unknown", set()
    allowed = os.sched_getaffinity(0) if hasattr(os, "sched_getaffinity") else set(range(os.cpu_count() or 1))
    try:
        cpu, phys, core = None, "0", None
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                key, _, val = line.partition(":")
                key, val = key.strip(), val.strip()
                if key == "processor":
                    cpu = int(val)
                elif key == "model name":
                    model = val
                elif key == "physical id":
                    phys = val
                elif key == "core id":
                    core = val
                elif not key and cpu is not None:
                    if cpu in allowed:
                        cores.add((phys, core if core is not None else str(cpu)))
                    cpu, phys, core = None, "0", None
        if cpu is not None and cpu in allowed:
            cores.add((phys, core if core is not None else str(cpu)))
    except OSError:
        pass
    return model, max(1, len(cores)) if cores else len(allowed)


# The oracle against the reference binary (oracle/_ref/HiFiLES_cpu, serial,
# no-BLAS dgemm, -O2), both on one core of the build host over the same 8^3 P=3
# Navier-Stokes TGV workload: measured by oracle/calibrate_ref.py, recorded in
# profiles/cpu_baseline_calibration.md.  The oracle is the faster of the two, so
# the baseline below overstates the reference's CPU speed.
CPU_CALIBRATION = {"host": "Intel(R) Xeon(R) Processor @ 2.70GHz", "workload": "8^3 hexes P=3 NS TGV, RK45, 1 core",
                   "reference_s_per_step": 0.9686, "oracle_s_per_step": 0.3449, "oracle_over_reference_speed": 2.808,
                   "source": "profiles/cpu_baseline_calibration.md (oracle/calibrate_ref.py)"}


def cpu_baseline(s, w, order, g):
    """The oracle (dense reference-order operators) on this host's cores, on the
    same mesh and state: one core, the median of 5 RK stages after a warm-up stage;
    all physical cores available (at most 16), the median of 5 RK45 steps after a
    warm-up step.  Wall-clock time (perf_counter)."""
    import statistics

    from oracle import pyoracle as po
    model, phys = cpu_model()
    threads = max(1, min(16, phys))

    def make(nthreads):
        if w["kind"] == "vortex":
            xv, c2v = g
            od = po.default_deck(equation=0, viscous=0, order=order, adv_type=3, dt=w["dt"], vcjh_scheme=1,
                                 dx_cyclic=VORTEX_L, dy_cyclic=VORTEX_L)
            o = po.Oracle(od, 1, xv, c2v, threads=nthreads)
            import frmesh
            o.set_state(frmesh.shu_vortex_state(o.get("pos_upts")))
            return o, "the same %d^2 P=%d Euler vortex workload" % (round(s.n_eles ** 0.5), order)
        if w["kind"] == "tet":
            import cases
            o = cases.oracle_of(g, threads=nthreads)
            o.set_state(cases.channel_tet_state(g, o.get("pos_upts")))
            return o, "the same 20,736-tet P=%d channel workload" % order
        xv = s.get("xv").reshape(-1, 3)
        c2v = s.get("c2v").reshape(-1, 8).astype(np.int32)
        od = po.default_deck(equation=0, viscous=1, order=order, adv_type=3, dt=w["dt"], tau=1.0, pen_fact=0.5,
                             gamma=1.4, prandtl=0.71, Mach_free_stream=0.1, nx_free_stream=1.0,
                             Re_free_stream=1600.0, L_free_stream=1.0, T_free_stream=300.0, vcjh_scheme=1,
                             dx_cyclic=TWO_PI, dy_cyclic=TWO_PI, dz_cyclic=TWO_PI)
        o = po.Oracle(od, 4, xv, c2v, threads=nthreads)
        o.set_state(tgv_state(o.get("pos_upts")))
        return o, "the same %d^3 P=%d Navier-Stokes workload" % (round(s.n_eles ** (1 / 3)), order)

    # all cores: whole RK45 steps
    o, what = make(threads)
    dof = o.nu * o.n_eles * o.nf
    o.zero_rk_register()
    o.run_steps(1)  # warm-up
    steps = []
    for _ in range(5):
        t0 = time.perf_counter()
        o.run_steps(1)
        steps.append(time.perf_counter() - t0)
    t_step = statistics.median(steps)
    del o
    # one core: single RK stages (residual + update), each a fifth of a step
    o, _ = make(1)
    o.zero_rk_register()
    stages = []
    for i in range(6):
        t0 = time.perf_counter()
        o.residual()
        o.advance(i % 5, 3)
        if i:  # the first is the warm-up
            stages.append(time.perf_counter() - t0)
    t_stage = statistics.median(stages)
    return {"value": round(dof * 5 / t_step / 1e9, 5), "unit": "GDOF-stage/s", "cores": threads, "kind": "port",
            "sample": "median of 5 RK45 steps (after 1 warm-up step) of %s, dense reference-order operators, "
                      "%d OpenMP threads, %.3g s per step" % (what, threads, t_step),
            "one_core": {"value": round(dof / t_stage / 1e9, 5), "unit": "GDOF-stage/s",
                         "sample": "median of 5 RK stages (after 1 warm-up stage), 1 thread, %.3g s per stage"
                                   % t_stage},
            "cpu": model, "physical_cores_available": phys, "calibration": CPU_CALIBRATION}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="tgv32")
    ap.add_argument("--n", type=int, default=0, help="hexes per direction (TGV workloads)")
    ap.add_argument("--order", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--per-point-metrics", action="store_true",
                    help="keep per-point metrics on the device even though the mesh is affine")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain copy kernel of MIB MiB (byte-counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the residual would shard by element partition with face-trace halos).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        nd = args.copy_calibration * (1 << 20) // 8
        ms = copy_ms(nd)
        print(json.dumps({"copy_bytes_read": nd * 8, "copy_bytes_written": nd * 8, "ms_per_copy": ms,
                          "GBs": 2 * nd * 8 / (ms * 1e-3) / 1e9}), flush=True)
        return
    w = WORKLOADS[args.workload]
    n = args.n if args.n else w.get("n", 0)
    order = args.order if args.order else w["order"]
    fp64 = measure(args, w, n, order, "fp64")
    mixed = measure(args, w, n, order, "mixed")
    s, r = fp64["solver"], fp64
    l2, mx = rel_errors(mixed["solver"].get("disu_upts"), s.get("disu_upts"))
    traffic = measured_traffic(args.workload, r["per_element"], "fp64") if (n, order) == (w.get("n", 0), w["order"]) \
        else None
    mtraffic = measured_traffic(args.workload, mixed["per_element"], "mixed") if traffic is not None else None
    kname = "fr_tet_mfma_kernel" if w["kind"] == "tet" else "fr_tensor_kernel"
    # the practical ceiling: a plain 1 GiB copy (one element per thread) timed now
    copy_gbs = 2 * COPY_DOUBLES * 8 / (copy_ms(COPY_DOUBLES) * 1e-3) / 1e9
    rec = {
        "metric": METRIC, "value": round(r["value"], 4), "unit": "GDOF-stage/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(r["wall"] * 1e3 / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": describe(w, n, order), "name": args.workload,
                   "n_eles": s.n_eles, "order": order, "dof": r["dof"], "stages_per_step": 5,
                   "metrics": "per element (affine elements)" if r["per_element"] else "per point"},
        "roofline": {"bound": "hbm", "achieved": round(r["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(r["achieved"] / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic["rk"] * s.n_eles) if traffic else None,
                     "traffic_source": TRAFFIC_SOURCE[args.workload] if traffic else None,
                     "kernel": "%s ADVANCE (flux, divergence, RK update, next traces)" % kname,
                     "bytes_per_launch": r["rk_b"] * s.n_eles, "bytes_per_element": r["rk_b"],
                     "avg_launch_ms": round(r["avg_rk_ms"], 5),
                     "copy_kernel_GBs": round(copy_gbs, 1), "frac_of_copy": round(r["achieved"] / copy_gbs, 4)},
        "stage": stage_block(r, s, traffic, w),
        "mixed": {"value": round(mixed["value"], 4), "unit": "GDOF-stage/s",
                  "speedup_vs_fp64": round(mixed["value"] / r["value"], 3),
                  "ms_per_step": round(mixed["wall"] * 1e3 / args.steps, 4),
                  "rel_l2_err_vs_fp64": [float("%.3e" % v) for v in l2],
                  "rel_max_err_vs_fp64": [float("%.3e" % v) for v in mx],
                  "err_after_steps": args.warmup + args.steps,
                  "roofline": {"achieved": round(mixed["achieved"], 1), "frac": round(mixed["achieved"] / HBM_PEAK_GBS, 4),
                               "traffic": round(mtraffic["rk"] * s.n_eles) if mtraffic else None,
                               "bytes_per_element": mixed["rk_b"], "avg_launch_ms": round(mixed["avg_rk_ms"], 5)},
                  "stage": stage_block(mixed, mixed["solver"], mtraffic, w)},
        "residual_L1": [float("%.6e" % v) for v in r["norms"]],
        "setup_s": round(r["setup"], 2),
    }
    if not args.per_point_metrics:
        pp = measure(args, w, n, order, "fp64", per_point=True)
        l2p, _ = rel_errors(pp["solver"].get("disu_upts"), s.get("disu_upts"))
        rec["per_point_fp64"] = {"value": round(pp["value"], 4), "unit": "GDOF-stage/s",
                                 "ms_per_step": round(pp["wall"] * 1e3 / args.steps, 4),
                                 "roofline": {"achieved": round(pp["achieved"], 1),
                                              "frac": round(pp["achieved"] / HBM_PEAK_GBS, 4),
                                              "bytes_per_element": pp["rk_b"],
                                              "avg_launch_ms": round(pp["avg_rk_ms"], 5)},
                                 "stage": stage_block(pp, pp["solver"], None, w),
                                 "max_rel_l2_diff_vs_per_element": float("%.3e" % l2p.max())}
        del pp
    if s.n_eles * s.nu <= GRAPH_MAX_POINTS:
        # hifr_run on a mesh this small replays one captured HIP graph per step
        # (the last step of a call runs eagerly); `value` above is the eager rate
        s.run(args.warmup)
        s.synchronize()
        t0 = time.perf_counter()
        s.run(args.steps)
        s.synchronize()
        tg = time.perf_counter() - t0
        rec["graph_replay"] = {"value": round(r["dof"] * 5 * args.steps / tg / 1e9, 4), "unit": "GDOF-stage/s",
                               "ms_per_step": round(tg * 1e3 / args.steps, 4)}
    if not args.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline(s, w, order, fp64["g"])
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec), flush=True)


def rel_errors(a, b):
    """Relative L2 and max error per conserved variable of a against b."""
    d = (a - b).reshape(-1, a.shape[-1])
    ref = b.reshape(-1, b.shape[-1])
    return np.sqrt((d ** 2).sum(0) / (ref ** 2).sum(0)), np.abs(d).max(0) / np.abs(ref).max(0)


def measure(args, w, n, order, precision, per_point=None):
    t_setup = time.perf_counter()
    per_point = args.per_point_metrics if per_point is None else per_point
    s, g = make_solver(w, n, order, args.steps, per_point, precision)
    t_setup = time.perf_counter() - t_setup
    s.run(args.warmup)
    s.synchronize()
    t0 = time.perf_counter()
    total_ms, kms, kl = s.run_timed(args.steps)
    s.synchronize()
    wall = time.perf_counter() - t0
    norms = s.norms(1)
    if not np.all(np.isfinite(norms)):
        raise SystemExit("non-finite residual (%s)" % precision)
    dof = s.nu * s.n_eles * s.nf
    per_element = s.metrics_per_element()
    nfaces = 4 if w["kind"] == "tet" else 2 * s.nd
    grad_b, rk_b = algorithmic_bytes(s.nu, s.nfp, nfaces, s.nd, per_element=per_element,
                                     mixed=precision == "mixed", viscous=w["kind"] != "vortex")
    avg_rk_ms = kms[2] / kl[2]
    avg_grad_ms = kms[1] / kl[1] if kl[1] else 0.0
    return {"solver": s, "g": g, "value": dof * 5 * args.steps / wall / 1e9, "wall": wall, "dof": dof,
            "norms": norms, "per_element": per_element, "grad_b": grad_b, "rk_b": rk_b, "avg_rk_ms": avg_rk_ms,
            "avg_grad_ms": avg_grad_ms, "achieved": rk_b * s.n_eles / (avg_rk_ms * 1e-3) / 1e9,
            "total_ms": total_ms, "setup": t_setup}


def stage_block(r, s, traffic, w):
    stage_ms = r["avg_rk_ms"] + r["avg_grad_ms"]
    stage_bytes = (r["grad_b"] + r["rk_b"]) * s.n_eles
    model = survey_model_bytes(s.nu, s.nfp, s.nd, les=w["kind"] == "tet") if w["kind"] != "vortex" else None
    grad_gbs = r["grad_b"] * s.n_eles / (r["avg_grad_ms"] * 1e-3) / 1e9 if r["avg_grad_ms"] else 0.0
    return {"avg_grad_kernel_ms": round(r["avg_grad_ms"], 5), "avg_rk_kernel_ms": round(r["avg_rk_ms"], 5),
            "grad_kernel_bytes_per_element": r["grad_b"],
            "grad_kernel_roofline": {"achieved": round(grad_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": round(grad_gbs / HBM_PEAK_GBS, 4)},
            "grad_kernel_traffic": round(traffic["grad"] * s.n_eles) if traffic else None,
            "algorithmic_bytes": stage_bytes,
            "achieved_GBs": round(stage_bytes / (stage_ms * 1e-3) / 1e9, 1),
            "three_phase_model_bytes_per_dof": round(model / s.nu / s.nf, 2) if model else None,
            "three_phase_model_GBs": round(model * s.n_eles / (stage_ms * 1e-3) / 1e9, 1) if model else None,
            "event_total_ms": round(r["total_ms"], 3)}


if __name__ == "__main__":
    main()
