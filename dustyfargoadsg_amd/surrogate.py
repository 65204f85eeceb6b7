"""The learned warm start: a small MLP that maps a disc's parameters to its
azimuthally averaged surface-density profile after a run, trained on runs of
the batched solver itself; the plain method then starts from the predicted
profile and judges it.

Network (dfargo_mlp_*, BF16 MFMA kernels with FP32 accumulation on the GPU):
4 inputs (log10 q, log10 alpha, h, flaring index, each scaled to [-1, 1]
over the sweep's ranges) -> 3 hidden layers of 256 GELU units -> 256
outputs, the ratio Sigma(r) / Sigma_init(r) on a fixed log grid of 256
radii over [0.4, 2.5].

The warm start multiplies the disc's initial density by the predicted ratio
(interpolated to the ring centres) and sets v_theta to the axisymmetric
equilibrium of that profile (gravity, pressure gradient and rotation); the
torque on the planet (UpdateLog's total torque) is then followed until it
stays within 5 % of the steady torque of the cold-started run.
"""
import argparse
import ctypes
import json
import math
import sys
import time

import numpy as np

from . import _capi
from .disc import Batch, Disc, DfargoError, _check
from .workloads import default_params, BC_WALL

R_GRID = 0.4 * (2.5 / 0.4) ** (np.arange(256) / 255.0)
RANGES = {"logq": (math.log10(1e-5), math.log10(3e-3)), "logalpha": (-4.0, -2.0),
          "h": (0.03, 0.08), "flaring": (0.0, 0.25)}
SIGMA_SLOPE = 0.5  # the sweep's discs share one density slope (not an input of the network)

_fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class MLP:
    """The surrogate network on one GPU."""

    def __init__(self, n_in=4, width=256, n_hidden=3, n_out=256, max_batch=4096, seed=1, device=0):
        self.lib = _capi.load()
        self.n_in, self.n_out = n_in, n_out
        h = ctypes.c_void_p()
        _check(self.lib, self.lib.dfargo_mlp_create(n_in, width, n_hidden, n_out, max_batch, seed, device,
                                                    ctypes.byref(h)))
        self.h = h
        self.n_params = self.lib.dfargo_mlp_param_count(h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.dfargo_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        out = np.zeros(self.n_params, dtype=np.float32)
        _check(self.lib, self.lib.dfargo_mlp_get_params(self.h, _fp(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        _check(self.lib, self.lib.dfargo_mlp_set_params(self.h, _fp(p)))

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], self.n_out), dtype=np.float32)
        _check(self.lib, self.lib.dfargo_mlp_forward(self.h, x.shape[0], _fp(x), _fp(y)))
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _check(self.lib, self.lib.dfargo_mlp_gradients(self.h, x.shape[0], _fp(x), _fp(t), _fp(loss), _fp(g)))
        return float(loss[0]), g

    def train(self, x, t, steps, lr=1e-3):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        _check(self.lib, self.lib.dfargo_mlp_train(self.h, x.shape[0], _fp(x), _fp(t), steps, lr, _fp(loss)))
        return float(loss[0])


# ---------------------------------------------------------------------------
# the sweep's discs

def sample_params(n, seed):
    """n parameter sets (q, alpha, h, flaring) from the batched sweep's ranges."""
    rng = np.random.default_rng(seed)
    return [dict(q=float(10 ** rng.uniform(*RANGES["logq"])), alpha=float(10 ** rng.uniform(*RANGES["logalpha"])),
                 h=float(rng.uniform(*RANGES["h"])), flaring=float(rng.uniform(*RANGES["flaring"])))
            for _ in range(n)]


def features(ps):
    """The network's inputs: each parameter scaled to [-1, 1] over its range."""
    def s(v, lo_hi):
        lo, hi = lo_hi
        return 2.0 * (v - lo) / (hi - lo) - 1.0
    return np.array([[s(math.log10(p["q"]), RANGES["logq"]), s(math.log10(p["alpha"]), RANGES["logalpha"]),
                      s(p["h"], RANGES["h"]), s(p["flaring"], RANGES["flaring"])] for p in ps], dtype=np.float32)


def disc_params(p, nrad, seed):
    """The isothermal planet disc of one parameter set (log grid over
    [0.4, 2.5], Nsec = 3 Nrad, planet on a fixed circular orbit at r = 1,
    smoothing 0.6 H)."""
    par = default_params(nrad=nrad, nsec=3 * nrad, rmin=0.4, rmax=2.5, log_grid=1, sigma0=6.3662e-4,
                         sigma_slope=SIGMA_SLOPE, aspect_ratio=p["h"], flaring_index=p["flaring"],
                         alpha_viscosity=p["alpha"], thickness_smoothing=0.6, forced_circular=1, add_noise=1,
                         noise_seed=seed, boundary=BC_WALL, fast_transport=1)
    planets = dict(dist=[1.0], mass=[p["q"]], acc=[0.0], feel_disk=[0], feel_others=[0])
    return par, planets


def _tables(par):
    lib = _capi.load()
    nr, ns = par.nrad, par.nsec
    radii, rmed, surf, az = np.zeros(nr + 1), np.zeros(nr), np.zeros(nr), np.zeros(ns)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    lib.dfargo_fill_polar_1d_arrays(ctypes.byref(par), dp(radii), dp(rmed), dp(surf), dp(az))
    return rmed


def profile_ratio(d, sigma_init_ring, rmed):
    """Sigma(r) / Sigma_init(r) of a disc, on the network's radius grid."""
    rho = d.state()["rho"][: d.nr * d.ns].reshape(d.nr, d.ns).mean(axis=1)
    return np.interp(np.log(R_GRID), np.log(rmed), rho / sigma_init_ring).astype(np.float32)


def warm_state(d, par, ratio_grid, rmed):
    """The disc's initial state with its density times the predicted ratio and
    v_theta in the axisymmetric equilibrium of the new profile."""
    st = d.state()
    nr, ns = d.nr, d.ns
    ratio = np.interp(np.log(rmed), np.log(R_GRID), ratio_grid)
    ratio = np.clip(ratio, 1e-3, None)
    rho = st["rho"][: nr * ns].reshape(nr, ns)
    rho *= ratio[:, None]
    sig = rho.mean(axis=1)
    h = par.aspect_ratio * rmed ** par.flaring_index
    cs2 = (h * rmed ** -0.5) ** 2
    P = sig * cs2
    dPdr = np.gradient(P, rmed)
    vt2 = rmed * (rmed ** -2.0 + dPdr / sig)
    vt = st["vt"][: nr * ns].reshape(nr, ns)
    vt[:] = np.sqrt(np.maximum(vt2, 0.0))[:, None] - par.omega_frame * rmed[:, None]
    return st


def total_torque(d):
    tq, _ = d.update_log()
    return float(tq[0, 0] + tq[0, 1])  # inner + outer disc torque on planet 0


def run_group(cases, orbits, nrad, prec=_capi.FP32, warm=None, log_torque=False, seed=0, log=None):
    """Runs the discs of `cases` (parameter dicts) for `orbits` orbits at r = 1
    in one batch; `warm[k]` (a ratio profile or None) warm-starts disc k.
    Returns the final ratio profiles and, with log_torque, the torque after
    every AlgoGas call of DT = 2 pi / 20."""
    DT = 2 * math.pi / 20
    discs, sig0, tables = [], [], []
    for k, p in enumerate(cases):
        par, pl = disc_params(p, nrad, seed + k)
        d = Disc(par, prec)
        d.initial_state(pl)
        rmed = _tables(par)
        sig0.append(par.sigma0 * rmed ** -par.sigma_slope)  # the initial profile without its noise
        tables.append(rmed)
        if warm is not None and warm[k] is not None:
            d.set_state(warm_state(d, par, warm[k], rmed))
        discs.append(d)
    b = Batch(discs)
    torques = [[] for _ in discs]
    for c in range(int(round(orbits * 20))):
        b.algogas(DT)
        if log and c % 200 == 199:
            log(f"{len(discs)} discs: {(c + 1) / 20:g} orbits")
        if log_torque:
            for k, d in enumerate(discs):
                torques[k].append(total_torque(d))
    prof = np.stack([profile_ratio(d, s, r) for d, s, r in zip(discs, sig0, tables)])
    b.close()
    for d in discs:
        d.close()
    return prof, np.array(torques)


def orbits_to_steady(torque, steady, tol=0.05, window=20):
    """Orbits (DT = 1/20 orbit) after which the one-orbit running mean of the
    torque stays within tol of `steady`."""
    t = np.asarray(torque)
    if len(t) < window:
        return float("nan")
    avg = np.convolve(t, np.ones(window) / window, mode="valid")
    ok = np.abs(avg - steady) <= tol * abs(steady)
    last_bad = np.nonzero(~ok)[0]
    first = 0 if len(last_bad) == 0 else last_bad[-1] + 1
    if first >= len(avg):
        return float("nan")
    return (first + window) / 20.0


def main(argv=None):
    ap = argparse.ArgumentParser(description="train and judge the surrogate warm start on one GPU")
    ap.add_argument("--ntrain", type=int, default=256)
    ap.add_argument("--nheld", type=int, default=32)
    ap.add_argument("--orbits", type=float, default=40.0)
    ap.add_argument("--nrad", type=int, default=128)
    ap.add_argument("--group", type=int, default=64)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    log = lambda *m: print("surrogate:", *m, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    train_p = sample_params(a.ntrain, a.seed)
    held_p = sample_params(a.nheld, a.seed + 1000)
    ys = []
    for g in range(0, a.ntrain, a.group):
        y, _ = run_group(train_p[g:g + a.group], a.orbits, a.nrad, seed=a.seed * 100000 + g, log=log)
        ys.append(y)
        log(f"training runs {g + len(y)}/{a.ntrain} ({time.perf_counter() - t0:.0f} s)")
    Y = np.concatenate(ys)
    X = features(train_p)
    t_data = time.perf_counter() - t0
    net = MLP(max_batch=max(a.ntrain, a.nheld), seed=a.seed)
    t1 = time.perf_counter()
    losses = []
    per = max(a.steps // 10, 1)
    for s in range(0, a.steps, per):
        losses.append(net.train(X, Y, min(per, a.steps - s), a.lr))
        log(f"step {s + per}: loss {losses[-1]:.3e}")
    t_train = time.perf_counter() - t1
    # judge on the held-out sets: cold start and warm start, same seeds
    pred = net.forward(features(held_p))
    cold_prof, cold_tq = run_group(held_p, a.orbits, a.nrad, log_torque=True, seed=777, log=log)
    log(f"cold held-out runs done ({time.perf_counter() - t0:.0f} s)")
    warm_prof, warm_tq = run_group(held_p, a.orbits, a.nrad, warm=list(pred), log_torque=True, seed=777, log=log)
    log(f"warm held-out runs done ({time.perf_counter() - t0:.0f} s)")
    nlast = max(int(0.1 * cold_tq.shape[1]), 1)
    res = []
    for k in range(a.nheld):
        steady = cold_tq[k, -nlast:].mean()
        oc = orbits_to_steady(cold_tq[k], steady)
        ow = orbits_to_steady(warm_tq[k], steady)
        res.append(dict(params=held_p[k], steady_torque=float(steady), orbits_cold=oc, orbits_warm=ow,
                        pred_l2=float(np.linalg.norm(pred[k] - cold_prof[k]) / np.linalg.norm(cold_prof[k])),
                        warm_final_l2=float(np.linalg.norm(warm_prof[k] - cold_prof[k]) /
                                            np.linalg.norm(cold_prof[k]))))
    oc = np.array([r["orbits_cold"] for r in res])
    ow = np.array([r["orbits_warm"] for r in res])
    both = np.isfinite(oc) & np.isfinite(ow)
    summary = dict(
        ntrain=a.ntrain, nheld=a.nheld, orbits=a.orbits, grid=f"{a.nrad}x{3 * a.nrad}", steps=a.steps,
        final_loss=losses[-1], loss_curve=losses, data_s=t_data, train_s=t_train,
        median_orbits_cold=float(np.median(oc[both])) if both.any() else None,
        median_orbits_warm=float(np.median(ow[both])) if both.any() else None,
        median_speedup=float(np.median(oc[both] / np.maximum(ow[both], 0.05))) if both.any() else None,
        reached_steady=[int(np.isfinite(oc).sum()), int(np.isfinite(ow).sum())],
        median_pred_l2=float(np.median([r["pred_l2"] for r in res])),
        median_warm_final_l2=float(np.median([r["warm_final_l2"] for r in res])),
        held_out=res)
    line = json.dumps(summary)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    print(line)


if __name__ == "__main__":
    main()
