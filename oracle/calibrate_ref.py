#!/usr/bin/env python3
"""Calibrates the oracle's CPU timing against the reference binary.

Both run the same Navier-Stokes Taylor-Green workload (hexes, P=3, Rusanov + LDG,
RK45) on one core: the reference build of oracle/Makefile.ref (serial, no-BLAS
dgemm, -O2) and the oracle (the C restatement, one OpenMP thread).  The reference
is timed around the whole process for 1 and for 1 + K steps; the difference,
divided by K, is its time per RK45 step without set-up and I/O.  The oracle is
timed per step after one warm-up step (median of K).

Needs /root/reference (cubature tables) and oracle/_ref/HiFiLES_cpu, so it runs
in the build container only; the numbers it prints are the calibration line that
bench.py's cpu_baseline quotes.

    python3 oracle/calibrate_ref.py [--n 8] [--order 3] [--steps 10]
"""
import argparse
import json
import os
import platform
import statistics
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

import frmesh  # noqa: E402
import make_goldens as mg  # noqa: E402
from oracle import pyoracle as po  # noqa: E402


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=8)
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--steps", type=int, default=10)
    args = ap.parse_args()
    if not os.path.exists(mg.REF_BIN):
        sys.exit("build the reference first: make -f oracle/Makefile.ref")
    n, P, K = args.n, args.order, args.steps
    c = dict(name="calibration", ele_type=4, n=(n, n, n), order=P, vcjh=1, viscous=1, tau=1.0, pen_fact=0.5,
             state="tgv", noise=1e-3, seed=1234, perturb=0.0, kind="rk45", dt=1e-4, adv_type=3)
    xv, conn_g = mg.build_case(c)
    c2v = frmesh.gambit_to_internal(4, conn_g)
    o = po.Oracle(mg.oracle_deck(c), 4, xv, c2v, threads=1)
    u0 = frmesh.tgv_state(o.get("pos_upts"), p0=1.0 / (1.4 * 0.1 ** 2))
    u0 = u0 * (1.0 + frmesh.seeded_perturbation(u0.shape, 1e-3, 1234))
    u0 = np.array([float("%.17g" % v) for v in u0.ravel()]).reshape(u0.shape)
    loc1d = o.get("loc_upts")[0, : P + 1]
    ref_t = {}
    for steps in (1, 1 + K):
        c["n_steps"] = steps
        t0 = time.perf_counter()
        u_ref, _ = mg.run_reference(c, xv, conn_g, u0, loc1d)
        ref_t[steps] = time.perf_counter() - t0
    ref_step = (ref_t[1 + K] - ref_t[1]) / K
    o.set_state(u0)
    o.zero_rk_register()
    o.run_steps(1)
    times = []
    for _ in range(K):
        t0 = time.perf_counter()
        o.run_steps(1)
        times.append(time.perf_counter() - t0)
    orc_step = statistics.median(times)
    # the oracle after 1 + K steps agrees with the reference's printed state
    rel = float(np.abs(o.get("disu_upts") - u_ref).max() / np.abs(u_ref).max())
    dof_stage = (P + 1) ** 3 * n ** 3 * 5 * 5
    rec = {"workload": "3-D Navier-Stokes TGV, %d^3 hexes, P=%d, RK45, 1 core" % (n, P), "cpu": cpu_model(),
           "reference_s_per_step": round(ref_step, 4), "oracle_s_per_step": round(orc_step, 4),
           "reference_GDOF_stage_s": round(dof_stage / ref_step / 1e9, 5),
           "oracle_GDOF_stage_s": round(dof_stage / orc_step / 1e9, 5),
           "oracle_over_reference_speed": round(ref_step / orc_step, 3), "state_rel_diff": float("%.2e" % rel)}
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
