"""Learned surrogate for parameter sweeps, with the evolver as fallback and as the measure of its error.

Inputs (8, sampled by Latin hypercube over RANGES): g, g', lambda, the rms
amplitude of the GRF Higgs fluctuations, the spectral slope of P(k) ~ |k|^slope,
the cut-off k_max as a fraction of pi/(2 dx), dx and dt.  Outputs: 64 time
points x 4 observables of an SM lattice evolved from that initial state
(mean |phi|^2/v^2, electric and magnetic energy fractions, mean 1x1 SU(2)
plaquette 1/2 tr), measured at steps round(k * steps / 63), k = 0..63.

Training data come from the batched evolver (gft_ensemble: 64 lattices per
launch, FP32-bulk).  The network (gft_mlp: 4 hidden layers of 256,
GELU, 4 heads, BF16 on the MFMA cores, trained on the GPU) maps the scaled
inputs to standardised outputs; each head is trained on its own Poisson(1)
bootstrap weights, so the heads agree where the data pin the map and spread
where they do not.  A sweep is screened by the network; a point whose head
spread is too large is rerun by the evolver, which also measures the error of
the points it accepts.

The spread cut is calibrated, not guessed (split conformal prediction): on a
held-out set of evolver runs (independent of the training runs and of the
sweep) the ratio error / spread is ranked, and q is its ceil((n+1)(1-alpha))-th
smallest value.  A point is accepted when q * spread <= eps.  For a sweep
point exchangeable with the calibration points, P(error > q * spread) <= alpha,
so at most a fraction alpha of the sweep is accepted with an error above eps:
the stated bound of an accepted point is eps (5% of the observable's range) at
confidence 1 - alpha (95%).  The evolver judges it on the whole sweep.

    python -m gftevolver_amd.surrogate [--train 4096] [--cal 512] [--sweep 512] [--L 64] [--steps 2000]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from ._abi import GFT_FORM_MOMENTUM_F32, make_desc

PARAMS = ("g", "gp", "lam", "rms", "slope", "kmax_frac", "dx", "dt")
RANGES = np.array([[0.5, 0.8], [0.25, 0.45], [0.08, 0.20], [0.05, 0.20], [-2.0, 0.0], [0.25, 1.0], [0.4, 0.6],
                   [0.02, 0.06]])
OBSERVABLES = ("phi_sq", "e_frac", "b_frac", "plaq_su2")
N_TIMES = 64


def latin_hypercube(n, seed):
    """n points in the box RANGES: each coordinate's n strata are each hit
    once, in an independent random order per coordinate, jittered inside the
    stratum."""
    rng = np.random.default_rng(seed)
    u = (np.stack([rng.permutation(n) for _ in PARAMS], axis=1) + rng.uniform(size=(n, len(PARAMS)))) / n
    return RANGES[:, 0] + u * (RANGES[:, 1] - RANGES[:, 0])


def scale_inputs(p):
    return (2.0 * (p - RANGES[:, 0]) / (RANGES[:, 1] - RANGES[:, 0]) - 1.0).astype(np.float32)


def time_points(steps, n_times=N_TIMES):
    return np.rint(np.arange(n_times) * steps / (n_times - 1)).astype(int)


def _desc(p, L):
    g, gp, lam, _, _, _, dx, dt = (float(v) for v in p)
    return make_desc(form=GFT_FORM_MOMENTUM_F32, n=(L, L, L), d=(dx, dx, dx), dt=dt, g=g, gp=gp, lam=lam)


def _init(ev, p, seed):
    """The point's GRF state, generated on the GPU (gft_init_grf)."""
    _, _, _, rms, slope, kfrac, dx, _ = (float(v) for v in p)
    ev.init_grf(seed=seed, rms=rms, slope=slope, kmax=kfrac * np.pi / (2 * dx))


def _observe(m):
    o = m.observe()
    w = m.wilson_loops()
    tot = o.energy[0]
    return (w.phi_sq, o.energy[5] / tot, o.energy[4] / tot, w.w11_su2)


def run_points(rt, points, L=64, steps=2000, members=64, seed=20261203, log=None):
    """Evolves every parameter point (FP32-bulk, batched ensembles of
    `members` lattices) and returns (Y [n, 64, 4], seconds of evolution,
    seconds in all).  Every point starts from the same white noise (GRF seed
    `seed`), filtered and scaled by its own spectrum and amplitude: with this
    common random number the observables are smooth functions of the
    parameters instead of carrying each run's own realisation noise."""
    tp = time_points(steps)
    n = len(points)
    Y = np.zeros((n, N_TIMES, len(OBSERVABLES)))
    t_evolve = 0.0
    t0 = time.perf_counter()
    batches = [range(b0, min(n, b0 + members)) for b0 in range(0, n, members)]
    for bi, idx in enumerate(batches):
        with rt.Ensemble([_desc(points[i], L) for i in idx]) as ens:
            for m, i in zip(ens.members, idx):
                _init(m, points[i], seed)
            for k in range(N_TIMES):
                if k:
                    t1 = time.perf_counter()
                    ens.evolve(int(tp[k] - tp[k - 1]))
                    t_evolve += time.perf_counter() - t1
                for j, m in enumerate(ens.members):
                    Y[idx[j], k] = _observe(m)
        if log and (bi % 8 == 7 or bi + 1 == len(batches)):
            log(f"  evolved points 0..{idx[-1]} ({time.perf_counter() - t0:.1f} s)")
    return Y, t_evolve, time.perf_counter() - t0


def init_params(layout, seed):
    """He-normal weights (std sqrt(2/fan_in)), zero biases, flat float32."""
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal(sh) * np.sqrt(2.0 / sh[1]) if len(sh) == 2 else np.zeros(sh) for _, sh in layout]
    return np.concatenate([q.ravel() for q in parts]).astype(np.float32)


class Surrogate:
    """The network plus its input scaling and output standardisation."""

    def __init__(self, rt, heads=4, width=256, hidden=4, max_batch=512):
        self.rt = rt
        self.heads = heads
        self.mlp = rt.Mlp(n_in=len(PARAMS), width=width, n_hidden=hidden, n_heads=heads,
                          n_out=N_TIMES * len(OBSERVABLES), max_batch=max_batch)

    def close(self):
        self.mlp.close()

    def fit(self, points, Y, epochs=1500, batch=128, lr=1e-3, seed=1):
        n = len(points)
        flat = Y.reshape(n, -1)
        self.mu = flat.mean(axis=0)
        # Per-observable range over data and time: the unit of the spread and error criteria.
        self.range = np.maximum(Y.max(axis=(0, 1)) - Y.min(axis=(0, 1)), 1e-300)
        # Standardise each output, with a floor of 1e-3 of its observable's range so that
        # outputs constant over the data (the energy fractions and plaquette at t = 0) stay
        # on the observable's scale.
        floor = np.tile(1e-3 * self.range, N_TIMES)
        self.sd = np.maximum(flat.std(axis=0), floor)
        rng = np.random.default_rng(seed)
        self.mlp.set_params(init_params(self.mlp.layout(), seed))
        X = scale_inputs(points)
        T = ((flat - self.mu) / self.sd).astype(np.float32)
        W = rng.poisson(1.0, size=(n, self.heads)).astype(np.float32)   # bootstrap weights per head
        hist = []
        # Three stages of decreasing learning rate (the Adam state carries over).
        for k, f in enumerate((1.0, 0.3, 0.1)):
            hist.append(self.mlp.train(X, T, W, batch=batch, epochs=epochs // 3, lr=lr * f, seed=seed + k))
        return np.concatenate(hist)

    def predict(self, points):
        """(mean [n, 64, 4], per-head [n, heads, 64, 4], spread [n])."""
        z = self.mlp.forward(scale_inputs(points)).astype(np.float64)     # [n, heads, 256]
        y = self.mu + self.sd * z
        y = y.reshape(len(points), self.heads, N_TIMES, len(OBSERVABLES))
        spread = ((y.max(axis=1) - y.min(axis=1)) / self.range).max(axis=(1, 2))
        return y.mean(axis=1), y, spread


def error(pred, truth, rng_):
    """Max over time and observables of |pred - truth| / observable range."""
    return (np.abs(pred - truth) / rng_).max(axis=(-2, -1))


def calibrate(sur, points, Y, eps=0.05, alpha=0.05):
    """Split-conformal spread cut from held-out evolver runs (points, Y):
    returns (threshold on the spread, q, calibration errors and spreads).  q is the
    ceil((n+1)(1-alpha))-th smallest error/spread ratio; accepting a point
    when its spread <= eps / q bounds P(accepted and error > eps) by alpha."""
    pred, _, spread = sur.predict(points)
    err = error(pred, Y, sur.range)
    ratio = np.sort(err / np.maximum(spread, 1e-12))
    n = len(ratio)
    k = int(np.ceil((n + 1) * (1.0 - alpha)))
    q = float(ratio[min(k, n) - 1])
    return eps / q, q, err, spread


def screen(sur, points, threshold):
    """Surrogate pass over a sweep: (prediction, spread, accepted mask, seconds)."""
    t0 = time.perf_counter()
    mean, _, spread = sur.predict(points)
    return mean, spread, spread <= threshold, time.perf_counter() - t0


def gain_study(rt, n_train=4096, n_sweep=512, L=64, steps=2000, epochs=1500, n_cal=512, eps=0.05, alpha=0.05,
               log=print):
    """Train on n_train evolver runs, calibrate the spread cut on n_cal
    held-out runs, screen an independent n_sweep-point Latin-hypercube sweep,
    and compare with evolving every sweep point (which also judges the error
    of every accepted point)."""
    P = latin_hypercube(n_train, seed=1)
    log(f"training data: {n_train} runs of SM {L}^3, {steps} steps")
    Y, te, tw = run_points(rt, P, L, steps, log=log)
    C = latin_hypercube(n_cal, seed=3)
    log(f"calibration data: {n_cal} runs")
    Yc, _, _ = run_points(rt, C, L, steps, log=log)
    sur = Surrogate(rt)
    try:
        t0 = time.perf_counter()
        hist = sur.fit(P, Y, epochs=epochs)
        t_fit = time.perf_counter() - t0
        log(f"trained in {t_fit:.1f} s; loss {hist[0]:.3e} -> {hist[-1]:.3e}")
        threshold, q, cal_err, cal_spread = calibrate(sur, C, Yc, eps, alpha)
        log(f"calibrated spread cut {threshold:.4f} (q = {q:.3f}; eps {eps}, alpha {alpha})")
        S = latin_hypercube(n_sweep, seed=2)
        log(f"sweep: {n_sweep} points, all evolved to judge the screen")
        truth, se, sw = run_points(rt, S, L, steps, log=log)
        pred, spread, ok, t_pred = screen(sur, S, threshold)
        train_pred, _, _ = sur.predict(P[:n_sweep])
    finally:
        sur.close()
    err = error(pred, truth, sur.range)
    train_err = error(train_pred, Y[:n_sweep], sur.range)
    per_point = sw / n_sweep
    screened_cost = t_pred + (~ok).sum() * per_point
    res = {
        "train_runs": n_train, "calibration_runs": n_cal, "sweep_points": n_sweep, "lattice": L, "steps": steps,
        "eps": eps, "alpha": alpha, "threshold": threshold, "conformal_q": q,
        "sweep_accepted_over_eps": int((ok & (err > eps)).sum()),
        "accepted_over_eps_fraction": float((err[ok] > eps).mean()) if ok.any() else None,
        "calibration_err_mean": float(cal_err.mean()),
        "evolver_seconds_per_point": per_point, "surrogate_seconds_per_sweep": t_pred, "train_seconds": t_fit,
        "data_seconds": tw, "accepted": int(ok.sum()), "rerun": int((~ok).sum()),
        "gain_cost": (n_sweep * per_point) / screened_cost,
        "err_accepted_max": float(err[ok].max()) if ok.any() else None,
        "err_accepted_p95": float(np.percentile(err[ok], 95)) if ok.any() else None,
        "err_accepted_mean": float(err[ok].mean()) if ok.any() else None,
        "err_rejected_mean": float(err[~ok].mean()) if (~ok).any() else None,
        "err_all_mean": float(err.mean()), "train_err_mean": float(train_err.mean()),
        "err_mean_by_observable": dict(zip(OBSERVABLES, (np.abs(pred - truth) / sur.range).max(axis=1).mean(axis=0).tolist())),
        "spread_median": float(np.median(spread)),
        "screened_sweep_err_max": float(np.where(ok, err, 0.0).max()),
        "loss_first_last": [float(hist[0]), float(hist[-1])],
        # per point: the head spread and the evolver-judged error of every sweep
        # point, and the calibration set's sorted error/spread ratios
        "sweep_spread": spread.tolist(), "sweep_err": err.tolist(),
        "calibration_ratio_sorted": np.sort(cal_err / np.maximum(cal_spread, 1e-12)).tolist(),
    }
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=4096)
    ap.add_argument("--cal", type=int, default=512)
    ap.add_argument("--sweep", type=int, default=512)
    ap.add_argument("--eps", type=float, default=0.05)
    ap.add_argument("--alpha", type=float, default=0.05)
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=1500)
    ap.add_argument("--out", type=str, default="")
    a = ap.parse_args(argv)
    from . import runtime as rt
    res = gain_study(rt, a.train, a.sweep, a.L, a.steps, a.epochs, a.cal, a.eps, a.alpha,
                     log=lambda *x: print(*x, flush=True))
    txt = json.dumps(res)
    print(txt, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
