"""Learned LES closure pipeline on the device: high-order Taylor-Green runs ->
filtered training samples -> network trained by the library's HIP kernels ->
coarse LES with the learned eddy viscosity (WALE fallback), judged by the
kinetic-energy dissipation rate of the filtered high-order run.

    python learned_sgs.py [--n-hi 32] [--p-hi 5] [--n-lo 64] [--re 5000] [--t-end 5]

Every step runs through libhifr.so (hifr.py): the solver, the sample kernel
(hifr_sgs_samples), the network's training kernels (hifr_mlp_*) and the LES flux
kernels with the network fused in (sgs_model HIFR_SGS_LEARNED).
"""
import argparse
import json
import math
import sys
import time

import numpy as np

import hifr

TWO_PI = 2.0 * math.pi
C_MAX = 2.0  # training targets are clipped to [0, C_MAX]


def tgv_deck(order, re, dt, les=False, learned=False, precision=0, filter_ratio=1.0):
    kw = dict(viscous=1, order=order, vcjh_scheme_hexa=1, dt=dt, adv_type=3, tau=1.0, pen_fact=0.5, gamma=1.4,
              prandtl=0.71, Mach_free_stream=0.1, nx_free_stream=1.0, Re_free_stream=float(re), L_free_stream=1.0,
              T_free_stream=300.0, fix_vis=1, dx_cyclic=TWO_PI, dy_cyclic=TWO_PI, dz_cyclic=TWO_PI,
              hifr_precision=precision)
    if les:
        kw.update(LES=1, SGS_model=1, filter_ratio=filter_ratio, filter_type=0, wall_model=0,
                  hifr_learned_sgs=1 if learned else 0)
    return hifr.deck_text(**kw)


def tgv_state(pos, gamma=1.4, mach=0.1):
    """Taylor-Green field (src/eles.cpp:560-568 with p0 = 1/(gamma Mach^2))."""
    x, y, z = pos[..., 0], pos[..., 1], pos[..., 2]
    p = 1.0 / (gamma * mach * mach) + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0) / 16.0
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    E = p / (gamma - 1.0) + 0.5 * (u * u + v * v)
    return np.stack([np.ones_like(x), u, v, np.zeros_like(x), E], axis=-1)


def gauss_weights3(order):
    w = np.polynomial.legendre.leggauss(order + 1)[1]
    return np.einsum("i,j,k->ijk", w, w, w).ravel()  # point index k + n j + n^2 i


def kinetic_energy(u, detj, order):
    """Volume average of rho |v|^2 / 2 over the periodic cube (Gauss rule of the
    solution points; detj per point (Nu, N))."""
    w3 = gauss_weights3(order)
    ke = 0.5 * (u[..., 1] ** 2 + u[..., 2] ** 2 + u[..., 3] ** 2) / u[..., 0]
    return float(np.einsum("j,je,je->", w3, detj, ke)) / TWO_PI ** 3


def stable_dt(n, order, mach=0.1):
    """RK45 step for the periodic TGV: 0.6 h / ((P+1)^2 (|v| + c)) with |v| = 1.
    (The 32^3 P=3 run is stable at a constant of 0.9; 48^3 P=4 at 1.05, not 2.1.)"""
    h = TWO_PI / n
    return 0.6 * h / ((order + 1) ** 2 * (1.0 + 1.0 / mach))


def run(solver, u0, dt, t_end, every, on_sample=None, log=None, sample_steps=()):
    """Steps to t_end; KE every `every` steps and at each step of sample_steps
    (and on_sample(solver, t, is_sample_step) at those stops)."""
    solver.set_state(u0)
    detj = solver.get("detjac_upts")
    order = int(round(solver.nu ** (1 / 3))) - 1
    steps = int(round(t_end / dt))
    samples = {int(k) for k in sample_steps if 0 < int(k) <= steps}
    stops = sorted(set(range(every, steps + 1, every)) | {steps} | samples)
    ts, kes = [0.0], [kinetic_energy(u0, detj, order)]
    if on_sample:
        on_sample(solver, 0.0, False)
    t0 = time.perf_counter()
    done = 0
    for stop in stops:
        solver.run(stop - done)
        done = stop
        u = solver.get("disu_upts")
        if not np.all(np.isfinite(u)):
            raise RuntimeError("non-finite state at t = %g" % (done * dt))
        ts.append(done * dt)
        kes.append(kinetic_energy(u, detj, order))
        if on_sample:
            on_sample(solver, done * dt, done in samples)
        if log:
            log("  t = %.3f  KE = %.6f  (%.1f s)" % (done * dt, kes[-1], time.perf_counter() - t0))
    return np.array(ts), np.array(kes), (time.perf_counter() - t0) / max(steps, 1)


def generate(n_hi, p_hi, re, t_end, every, p_lo=2, sample_times=(), log=None):
    """High-order run: KE history of the filtered field (degree p_lo) and training
    samples at the steps nearest to the sample times (one snapshot per time)."""
    dt = stable_dt(n_hi, p_hi)
    s = hifr.Solver.box(tgv_deck(p_hi, re, dt), hifr.HEX, (n_hi,) * 3, (TWO_PI,) * 3, (0.0,) * 3)
    u0 = tgv_state(s.get("pos_upts"))
    h = TWO_PI / n_hi
    detj_lo = np.full(((p_lo + 1) ** 3, s.n_eles), (h / 2) ** 3)
    ke_f, X, C, CW, tagged = [], [], [], [], []

    def on_sample(solver, t, is_sample):
        x, ce, cw, ulo = solver.sgs_samples(p_lo, 1.0, want_state=True)
        ke_f.append(kinetic_energy(ulo, detj_lo, p_lo))
        if is_sample:
            X.append(x)
            C.append(ce)
            CW.append(cw)
            tagged.append(t)

    sample_steps = [int(round(ts / dt)) for ts in sample_times]
    ts, ke, ms = run(s, u0, dt, t_end, every, on_sample, log, sample_steps)
    return {"t": ts, "ke": ke, "ke_filtered": np.array(ke_f), "dt": dt, "ms_per_step": ms * 1e3,
            "samples": (X, C, CW, tagged)}


def normalization(x):
    mean = x.mean(axis=0).astype(np.float32)
    scale = x.std(axis=0).astype(np.float32)
    scale[scale <= 0] = 1.0
    return mean, scale, x.min(axis=0).astype(np.float32), x.max(axis=0).astype(np.float32)


def train(x, c, epochs=20, batch=4096, lr=2e-3, seed=0, log=None):
    """Network fitted on the device to the clipped exact coefficients."""
    target = np.clip(c, 0.0, C_MAX)
    out_scale = float(max(target.mean(), 1e-6))
    mean, scale, lo, hi = normalization(x)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(c))
    xs = ((x - mean) / scale).astype(np.float32)[perm]
    ys = (target / out_scale).astype(np.float32)[perm]
    m = hifr.MLP(seed=seed)
    m.set_normalization(mean, scale, lo, hi, out_scale)
    m.set_data(xs, ys)
    hist = []
    left = epochs
    while left > 0:
        k = min(left, 50)
        hist.extend(m.train(k, batch, lr).tolist())
        left -= k
    if log:
        log("  training loss per epoch: %s" % " ".join("%.4g" % v for v in hist))
    return m, (mean, scale, lo, hi, out_scale), hist


def predict_c(m, norm, x):
    mean, scale, _, _, out_scale = norm
    return m.predict(((x - mean) / scale).astype(np.float32)) * out_scale


def apriori(m, norm, x, c, cw):
    """Relative RMS error of the coefficient against the clipped exact one, for
    the network and for WALE, and their correlations."""
    target = np.clip(c, 0.0, C_MAX)
    cp = predict_c(m, norm, x)
    rel = lambda a: float(np.sqrt(np.mean((a - target) ** 2) / np.mean(target ** 2)))
    corr = lambda a: float(np.corrcoef(a, target)[0, 1])
    return {"rel_rms_learned": rel(cp), "rel_rms_wale": rel(cw), "corr_learned": corr(cp), "corr_wale": corr(cw)}


def deploy(n, p, re, t_end, every, model, network=None, precision=0, log=None):
    """Coarse run: model 'none', 'wale' (the learned kernels without a network) or
    'learned'.  Returns (t, KE, ms per step)."""
    dt = stable_dt(n, p)
    les = model != "none"
    s = hifr.Solver.box(tgv_deck(p, re, dt, les=les, learned=les, precision=precision), hifr.HEX, (n,) * 3,
                        (TWO_PI,) * 3, (0.0,) * 3)
    if model == "learned":
        s.set_sgs_network(network)
    u0 = tgv_state(s.get("pos_upts"))
    ts, ke, ms = run(s, u0, dt, t_end, every, None, log)
    return ts, ke, ms * 1e3


def time_models(n, p, re, steps, log=None):
    """Device time per RK45 step of the coarse LES with no model, WALE and the
    network (random weights with a normalisation whose range admits every input,
    so that every point evaluates the network: its full cost), after 5 untimed
    steps.  Returns {model: ms per step}."""
    import time
    dt = stable_dt(n, p)
    out = {}
    for model in ("none", "wale", "learned"):
        les = model != "none"
        s = hifr.Solver.box(tgv_deck(p, re, dt, les=les, learned=les), hifr.HEX, (n,) * 3, (TWO_PI,) * 3,
                            (0.0,) * 3)
        if model == "learned":
            m = hifr.MLP(seed=3)
            m.set_normalization(np.zeros(10, np.float32), np.ones(10, np.float32), np.full(10, -1e30, np.float32),
                                np.full(10, 1e30, np.float32), 1.0)
            s.set_sgs_network(m)
        s.set_state(tgv_state(s.get("pos_upts")))
        s.run(5)
        s.synchronize()
        t0 = time.perf_counter()
        s.run(steps)
        s.synchronize()
        out[model] = (time.perf_counter() - t0) / steps * 1e3
        if log:
            log("%-8s %.3f ms per RK45 step" % (model, out[model]))
    return out


def dissipation(t, ke):
    return -np.gradient(ke, t)


def eps_error(t_ref, eps_ref, t, eps):
    """Relative L2 error of a dissipation-rate history against the reference one
    (interpolated to the reference times)."""
    e = np.interp(t_ref, t, eps)
    return float(np.sqrt(np.mean((e - eps_ref) ** 2) / np.mean(eps_ref ** 2)))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-hi", type=int, default=32)
    ap.add_argument("--p-hi", type=int, default=5)
    ap.add_argument("--n-lo", type=int, default=64)
    ap.add_argument("--p-lo", type=int, default=2)
    ap.add_argument("--re", type=float, default=5000.0)
    ap.add_argument("--t-end", type=float, default=5.0)
    ap.add_argument("--samples-at", type=str, default="2,3,4,5")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--out", type=str, default="")
    ap.add_argument("--time-steps", type=int, default=0,
                    help="only time this many RK45 steps of each coarse model and exit")
    args = ap.parse_args(argv)
    log = lambda msg: print(msg, flush=True)
    if args.time_steps > 0:
        log("coarse LES step time: %d^3 hexes, P=%d, Re %g" % (args.n_lo, args.p_lo, args.re))
        return time_models(args.n_lo, args.p_lo, args.re, args.time_steps, log)
    times = [float(v) for v in args.samples_at.split(",")]
    log("high-order run: %d^3 hexes, P=%d, Re %g" % (args.n_hi, args.p_hi, args.re))
    gen = generate(args.n_hi, args.p_hi, args.re, args.t_end, 50, args.p_lo, times, log)
    X, C, CW, tagged = gen["samples"]
    x, c, cw = np.concatenate(X[:-1]), np.concatenate(C[:-1]), np.concatenate(CW[:-1])
    log("samples: %d for training (t = %s), %d held out (t = %g)" % (len(c), tagged[:-1], len(C[-1]), tagged[-1]))
    m, norm, hist = train(x, c, epochs=args.epochs, log=log)
    ap_stats = apriori(m, norm, X[-1], C[-1], CW[-1])
    log("a priori (held-out snapshot): %s" % json.dumps(ap_stats))
    eps_ref = dissipation(gen["t"], gen["ke_filtered"])
    out = {"high_order": {"n": args.n_hi, "p": args.p_hi, "ms_per_step": gen["ms_per_step"], "t": gen["t"].tolist(),
                          "ke": gen["ke"].tolist(), "ke_filtered": gen["ke_filtered"].tolist()},
           "train_loss": hist, "apriori": ap_stats, "les": {}}
    for model in ("none", "wale", "learned"):
        log("LES %s: %d^3 hexes, P=%d" % (model, args.n_lo, args.p_lo))
        t, ke, ms = deploy(args.n_lo, args.p_lo, args.re, args.t_end, 50, model, m, log=log)
        err = eps_error(gen["t"], eps_ref, t, dissipation(t, ke))
        out["les"][model] = {"t": t.tolist(), "ke": ke.tolist(), "ms_per_step": ms, "eps_rel_err": err}
        log("  dissipation-rate error vs filtered high-order run: %.4f, %.3f ms per step" % (err, ms))
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f)
    return out


if __name__ == "__main__":
    main(sys.argv[1:])
