#!/usr/bin/env python3
"""Hydro sub-step throughput on one MI355X.

A step is one hydro sub-step of AlgoGas (one pass of the reference's
`while (dtemp < DT)` body) over the whole disc; the metric is cell-updates
per second with the state resident in HBM.  The default workload is the
adiabatic self-gravitating disc at 512 x 1536 (energy equation, beta cooling
10, alpha 1e-3, SG by 1024 x 1536 FFT), timed with FP32 storage (the
headline `value`) and with FP64 storage beside it.  `--workload gas_dust`
runs the 1024 x 3072 gas disc with 10^6 Lagrangian particles (drag
feedback on), whose sub-step also moves and deposits the particles.

Prints one JSON line.  Per-kernel times come from HIP events on the
library's stream in a second pass over the same sub-steps; the roofline
entry is the kernel with the largest share of the step time.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def kernel_bytes_per_cell(name, w, energy, sg):
    """Compulsory HBM bytes per cell of one launch of each kernel: every
    field it must read once and write once (tables and scalars count 0)."""
    ns = 4 if energy else 3                  # Sigma, vr, vt (+ e)
    nq = 6 if energy else 5                  # transported quantities
    table = {
        # CFL, force/mass partial sums, crash scan (+ sound speed: e in, cs out)
        "cfl_ring": (3 + (2 if energy else 0)) * w,
        # SubStep1 + viscosity + SubStep2: state (+ g_r, g_theta) in, new velocities (+ e) out
        "source": (ns + (2 if sg else 0)) * w + (2 + (1 if energy else 0)) * w,
        "energy": 4 * w + 1 * w,                         # rho, vr, vt, e in; e out
        "transport_rad": ns * w + nq * w,                # state in, radially transported quantities out
        "transport_azi": (nq + 1) * w + nq * w,          # quantities + vt in, shifted quantities out
        "finish": nq * w + ns * w,                       # momenta in, state out
        "sg_rows_fwd": w + 2 * w,                        # Sigma -> complex row spectra
        # Z columns in, the two half kernel spectra ((Ns/2+1) x 2 Nrad complex
        # each, 4 words per cell), W columns out
        "sg_cols": 2 * w + 4 * w + 2 * w,
        "sg_rows_inv": 2 * w + w + 2 * w,                # W row, Sigma -> g_r, g_theta
        "control": 0,
    }
    return table.get(name, 0)


def kernel_bytes_per_launch(name, w, energy, sg, cells, npart):
    """Compulsory HBM bytes of one launch: the per-cell figure times the
    cells, and for the particle kernels the particle arrays each must read
    and write once (FP64 particles, int32 cells and buckets)."""
    if name.startswith("dust_"):
        per_particle = {
            # r, th, vr, vth, size in; r, th, vr, vth, l, St, drag x2, t_s,
            # |dv|^2, the CIC/TSC weights (12-18 doubles) out; cell and slot
            "dust_move": 5 * 8 + 10 * 8 + 12 * 8 + 4 + 4,
            "dust_scan": 0,
            "dust_scatter": 4 + 4 + 4,             # cell, slot in; bucket entry out
            "dust_sort": 0,
            # bucket entry, r, th, cell, drag x2, four weights per stencil (+ t_s,
            # |dv|^2 with the energy equation)
            "dust_deposit": 8 + 4 + 4 + 16 + 8 * 8 + (16 if energy else 0),
            # bucket entry in; seven FP64 fields and the id read and written
            "dust_reorder": 8 + 2 * (7 * 8 + 4),
        }[name]
        per_cell = {
            "dust_move": 0,
            "dust_scan": 4 + 4,                    # counts in, ring-local starts out
            "dust_scatter": 4,                     # ring-local start of the particle's cell
            "dust_sort": 4 + 4 + 4,                # local start, count in; start out
            # starts, counts, rho in; two feedback grids (+ FBedot) out
            "dust_deposit": 8 + w + 2 * w + (w if energy else 0),
            "dust_reorder": 0,
        }[name]
        return per_particle * npart + per_cell * cells
    return kernel_bytes_per_cell(name, w, energy, sg) * cells


def step_bytes_per_cell(w, energy, sg):
    """Algorithmic bytes per cell per sub-step: each state field read and
    written once, plus the two kernel spectra of self-gravity (2 x 2Nr x
    (Ns/2+1) complex values, ~4 words per cell)."""
    ns = 4 if energy else 3
    return 2 * ns * w + (4 * w if sg else 0)


def run_gpu(params, planets, precision, steps, warmup, prof_steps, dust=None):
    from dustyfargoadsg_amd.disc import Disc, init_dust_system
    d = Disc(params, precision)
    d.initial_state(planets)
    if dust is not None:
        d.attach_dust(dust, init_dust_system(params, dust), 2 * math.pi / 20)
    DT = 1.0e3  # an AlgoGas call much longer than the CFL step
    d.run_substeps(DT, warmup)
    t0 = time.perf_counter()
    d.run_substeps(DT, steps)  # returns after the stream has drained
    t1 = time.perf_counter()
    st = d.status()
    assert st.last_call_substeps == steps, (st.last_call_substeps, steps)
    d.set_profiling(True)
    d.run_substeps(DT, prof_steps)
    kt = d.kernel_times()
    d.set_profiling(False)
    d.close()
    return t1 - t0, kt


def dust_fluid_timing(params, planets, mode, steps, warmup, prof_steps=40):
    """The workload with the dust fluid (FP64 storage; "sfta": the
    short-friction-time approximation, "full": the dust's momentum equations
    with viscosity and diffusion; 1 cm grains) beside the gas alone: ms per
    sub-step of each and the fluid's kernels (HIP events, one launch at a
    time)."""
    from dustyfargoadsg_amd import _capi
    from dustyfargoadsg_amd.disc import Disc
    DT = 1.0e3
    out = {}
    for name in ("gas", mode):
        d = Disc(params, _capi.FP64)
        if name != "gas":
            full = mode == "full"
            d.attach_dust_fluid(_capi.dust_fluid_default_params(
                sfta=0 if full else 1, sizepart=1e-2, dviscosity=1e-6 if full else 0.0,
                dust_diffusion=1 if full else 0))
        d.initial_state(planets)
        d.run_substeps(DT, warmup)
        t0 = time.perf_counter()
        d.run_substeps(DT, steps)
        t1 = time.perf_counter()
        d.set_profiling(True)
        d.run_substeps(DT, prof_steps)
        kt = d.kernel_times()
        d.close()
        out[name] = dict(ms_per_step=(t1 - t0) / steps * 1e3,
                         kernels_us={k: round(v[0] / v[1] * 1e3, 2) for k, v in kt.items() if v[1]})
    cells = params.nrad * params.nsec
    return {"mode": mode, "cells": cells, "steps": steps,
            "gas_ms_per_step": out["gas"]["ms_per_step"], "with_fluid_ms_per_step": out[mode]["ms_per_step"],
            "fluid_share": 1.0 - out["gas"]["ms_per_step"] / out[mode]["ms_per_step"],
            "fluid_kernels_us": {k: v for k, v in out[mode]["kernels_us"].items() if k.startswith("df_")}}


def batched_sweep(n, seed, calls, warmup_calls=1, tune=True):
    """The batched parameter sweep: n isothermal FP32 planet discs of uneven
    grids and parameters (workloads.batch_discs), advanced `calls` AlgoGas
    calls of DT = 2 pi / 20 each (1/20 orbit at r = 1) through one batch
    (one graph launch per sub-step for all discs), against a second set of
    the same discs advanced one disc after another (each disc's own graph).
    Launch shapes from the autotuner (one tuning per grid shape).  Returns
    discs x orbits per hour for both and the sub-steps taken."""
    from dustyfargoadsg_amd import _capi
    from dustyfargoadsg_amd.disc import Batch, Disc
    from dustyfargoadsg_amd.workloads import batch_discs
    DT = 2 * math.pi / 20
    cases = batch_discs(n=n, seed=seed)
    sets = []
    t0 = time.perf_counter()
    for _ in range(2):
        ds = []
        for p, pl in cases:
            d = Disc(p, _capi.FP32)
            d.initial_state(pl)
            ds.append(d)
        sets.append(ds)
    t_create = time.perf_counter() - t0
    tunings = {}
    t0 = time.perf_counter()
    if tune:
        for ds in sets:
            for d in ds:
                tunings[f"{d.nr}x{d.ns}"] = d.autotune(substeps=4)
    t_tune = time.perf_counter() - t0
    batched, seq = sets
    b = Batch(batched)
    for _ in range(warmup_calls):
        b.algogas(DT)
        for d in seq:
            d.algogas(DT)

    def timed(step, discs):
        sub0 = [d.status().substeps_total for d in discs]
        t0 = time.perf_counter()
        for _ in range(calls):
            step()
        el = time.perf_counter() - t0
        steps = [d.status().substeps_total - s0 for d, s0 in zip(discs, sub0)]
        return el, sum(steps), sum(k * d.nr * d.ns for k, d in zip(steps, discs))

    # the default dispatch and the sequential runs over the same calls of
    # the same discs (so the same sub-steps), then the one-graph dispatch
    # over the calls after them
    b.set_mode(Batch.STREAMS)
    t_b, sub_b, cu_b = timed(lambda: b.algogas(DT), batched)
    t_s, sub_s, cu_s = timed(lambda: [d.algogas(DT) for d in seq], seq)
    b.set_mode(Batch.GRAPH)
    t_g, sub_g, cu_g = timed(lambda: b.algogas(DT), batched)
    orbits = calls * DT / (2 * math.pi)
    cells = sum(p.nrad * p.nsec for p, _ in cases)
    b.close()
    for ds in sets:
        for d in ds:
            d.close()
    return {
        "workload": f"{n} isothermal FP32 planet discs, Nrad in {{128, 256, 512}}, Nsec = 3 Nrad, seed {seed}",
        "grids": {k: sum(1 for p, _ in cases if f"{p.nrad}x{p.nsec}" == k) for k in
                  sorted({f"{p.nrad}x{p.nsec}" for p, _ in cases})},
        "calls": calls, "orbits_per_disc": orbits,
        "batched_discs_orbits_per_hour": n * orbits / t_b * 3600.0,
        "sequential_discs_orbits_per_hour": n * orbits / t_s * 3600.0,
        "speedup": t_s / t_b,
        "batched_s": t_b, "sequential_s": t_s,
        "substeps": [sub_b, sub_s],
        "cell_updates_per_s": [cu_b / t_b, cu_s / t_s],
        "one_graph_mode": {"discs_orbits_per_hour": n * orbits / t_g * 3600.0, "s": t_g, "substeps": sub_g,
                           "cell_updates_per_s": cu_g / t_g},
        "cells_total": cells,
        "tuning": tunings, "create_s": t_create, "tune_s": t_tune,
    }


def precision_error(params, planets, orbits, control=True, checkpoints=None, noise=(1e-14, 2.0 ** -24),
                    reinject=()):
    """FP32 against FP64 storage from the same initial disc after `orbits`
    orbits at r = 1 (AlgoGas calls of DT = 2 pi / 20): relative L-infinity
    and L2 errors per field (Sigma, v_theta, energy on their own scale; v_r
    on the v_theta scale), the relative L-infinity error of the ring-mean
    Sigma(r) and of the disc mass, and the largest Sigma error in each of 16
    radial bands.  With `control`, FP64 runs whose initial Sigma carries
    independent relative noise of each amplitude in `noise` (uniform in
    [-eps, eps] per cell, seeded) give the same figures for FP64 against
    itself: 1e-14 shows how fast the flow amplifies perturbations, 2^-24
    what one FP32 rounding of the initial state alone grows into.  Where the
    flow amplifies perturbations (gravitoturbulence, unstable gap edges),
    that is the floor any pointwise comparison has.  Each amplitude in
    `reinject` adds an FP64 run that receives fresh noise of that amplitude
    in Sigma, v_r, v_theta (and e) after every AlgoGas call: perturbations
    of FP32 size that keep coming, as FP32 storage brings them, but no
    rounding inside the sub-steps."""
    import numpy as np
    from dustyfargoadsg_amd import _capi
    from dustyfargoadsg_amd.disc import Disc
    DT = 2 * math.pi / 20
    nr, ns = params.nrad, params.nsec
    n = nr * ns

    marks = sorted(set(int(round(o * 20)) for o in (checkpoints or []) if 0 < o < orbits))

    def run(prec, eps, again=0.0):
        d = Disc(params, prec)
        st = d.initial_state(planets)
        rng = np.random.default_rng(12345)
        if eps:
            st["rho"][:n] *= 1.0 + eps * rng.uniform(-1.0, 1.0, n)
            d.set_state(dict(rho=st["rho"], vr=st["vr"], vt=st["vt"], energy=st["energy"]))
        saved = {}
        for c in range(int(round(orbits * 20))):
            d.algogas(DT)
            if again:
                s2 = d.state()
                s2["rho"][:n] *= 1.0 + again * rng.uniform(-1.0, 1.0, n)
                s2["vt"][:n] *= 1.0 + again * rng.uniform(-1.0, 1.0, n)
                s2["vr"][:n] += again * np.abs(s2["vt"][:n]) * rng.uniform(-1.0, 1.0, n)
                if params.energy_equation:
                    s2["energy"][:n] *= 1.0 + again * rng.uniform(-1.0, 1.0, n)
                d.set_state(dict(rho=s2["rho"], vr=s2["vr"], vt=s2["vt"], energy=s2["energy"]))
            if c + 1 in marks:
                saved[c + 1] = d.state()
                saved[c + 1]["substeps"] = d.status().substeps_total
            if c % 200 == 199:
                print(f"precision run {'fp64' if prec == _capi.FP64 else 'fp32'}{f' noise {eps:g}' if eps else ''}:"
                      f" {(c + 1) / 20:g} orbits", file=sys.stderr, flush=True)
        out = d.state()
        out["substeps"] = d.status().substeps_total
        out["saved"] = saved
        d.close()
        return out

    def compare(a, b):
        def linf(k, scale=None):
            ref = b[scale or k][:n]
            return float(np.abs(a[k][:n] - b[k][:n]).max() / np.abs(ref).max())

        def l2(k, scale=None):
            ref = b[scale or k][:n]
            return float(np.linalg.norm(a[k][:n] - b[k][:n]) / np.linalg.norm(ref))
        pa, pb = a["rho"][:n].reshape(nr, ns).mean(axis=1), b["rho"][:n].reshape(nr, ns).mean(axis=1)
        out = {"sigma_linf": linf("rho"), "sigma_l2": l2("rho"),
               "sigma_ring_mean_linf": float(np.abs(pa - pb).max() / np.abs(pb).max()),
               "mass": float(abs(a["rho"][:n].sum() / b["rho"][:n].sum() - 1.0)),
               "vtheta_linf": linf("vt"), "vr_linf_on_vtheta": linf("vr", "vt"), "vr_l2_on_vtheta": l2("vr", "vt")}
        if params.energy_equation:
            out["energy_linf"] = linf("energy")
            out["energy_l2"] = l2("energy")
        band = np.abs(a["rho"][:n] - b["rho"][:n]).reshape(nr, ns).max(axis=1) / np.abs(b["rho"][:n]).max()
        out["sigma_linf_by_band"] = [float(x.max()) for x in np.array_split(band, 16)]
        out["substeps"] = [int(a["substeps"]), int(b["substeps"])]
        qa, qb = a["planets"], b["planets"]
        if qb.nb > 0:
            out["planet_position"] = max(max(abs(qa.x[k] - qb.x[k]), abs(qa.y[k] - qb.y[k])) for k in range(qb.nb))
            out["planet_mass"] = max(abs(qa.mass[k] / qb.mass[k] - 1.0) for k in range(qb.nb))
        return out

    ref = run(_capi.FP64, 0.0)
    f32 = run(_capi.FP32, 0.0)
    res = {"orbits": orbits, "fp32_vs_fp64": compare(f32, ref)}
    for c in marks:
        res[f"fp32_vs_fp64_at_{c / 20:g}_orbits"] = compare(f32["saved"][c], ref["saved"][c])
    if control:
        for eps in noise:
            res[f"fp64_noise_{eps:.3g}_vs_fp64"] = compare(run(_capi.FP64, eps), ref)
    for eps in reinject:
        res[f"fp64_noise_{eps:.3g}_every_call_vs_fp64"] = compare(run(_capi.FP64, 0.0, eps), ref)
    return res


def cpu_baseline(params, planets, budget_s, dust=None):
    """The FP64 oracle (a single-threaded C restatement of the sub-step) on
    the host, one sub-step per AlgoGas call, for about budget_s seconds."""
    from tests.oracle_lib import Oracle
    o = Oracle(params)
    o.initial_state(planets)
    if dust is not None:
        from dustyfargoadsg_amd.disc import init_dust_system
        o.dust_attach(dust, init_dust_system(params, dust), 2 * math.pi / 20)
    n, t0 = 0, time.perf_counter()
    while True:
        o.algogas(1e-7)  # shorter than the CFL step: one sub-step per call
        n += 1
        el = time.perf_counter() - t0
        if el > budget_s and n >= 2:
            break
    return n * params.nrad * params.nsec / el, n, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="energy_sg")
    ap.add_argument("--nrad", type=int, default=None)
    ap.add_argument("--nsec", type=int, default=None)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device three times (the known-size "
                         "reference for scaling FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--error-orbits", type=float, default=1.0,
                    help="orbits over which the FP32 error against FP64 is measured (0: skip)")
    ap.add_argument("--target-steps", type=int, default=100,
                    help="sub-steps of the 1024x3072 FP32 gas+SG disc timed beside the default "
                         "workload (0: skip)")
    ap.add_argument("--batched", type=int, default=64,
                    help="discs of the batched sweep timed beside the default workload (0: skip)")
    ap.add_argument("--batched-calls", type=int, default=2)
    ap.add_argument("--reinject", action="store_true",
                    help="with --precision-only, add the FP64 run that receives FP32-size noise every call")
    ap.add_argument("--no-controls", action="store_true",
                    help="with --precision-only, skip the perturbed FP64 control runs")
    ap.add_argument("--dust-fluid", choices=("sfta", "full"), default=None,
                    help="time the workload with the dust fluid (FP64) beside the gas alone, print that, return")
    ap.add_argument("--precision-only", type=float, default=0.0,
                    help="only run the FP32-vs-FP64 comparison of the workload over this many orbits "
                         "(with checkpoints every tenth of it and the perturbed FP64 control) and print it")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs one disc on one GPU; multi-GPU runs are not "
              "implemented (the ring-slab sharding is a design only)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = 1 << 26  # 64 Mi floats = 256 MiB
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(f"copied 3 x {n * 4} bytes")
        return
    from dustyfargoadsg_amd import _capi
    from dustyfargoadsg_amd.workloads import WORKLOADS
    kw = {}
    if args.nrad:
        kw["nrad"] = args.nrad
    if args.nsec:
        kw["nsec"] = args.nsec
    params, planets = WORKLOADS[args.workload](**kw)
    from dustyfargoadsg_amd.workloads import DUST_WORKLOADS
    dust = DUST_WORKLOADS[args.workload](params) if args.workload in DUST_WORKLOADS else None
    npart = dust.nbpart if dust is not None else 0
    if args.dust_fluid:
        print(json.dumps({"workload": f"{args.workload} {params.nrad}x{params.nsec}",
                          **dust_fluid_timing(params, planets, args.dust_fluid, args.steps, args.warmup)}))
        return
    if args.precision_only > 0:
        o = args.precision_only
        print(json.dumps({"workload": f"{args.workload} {params.nrad}x{params.nsec}",
                          **precision_error(params, planets, o, control=not args.no_controls,
                                            reinject=(2.0 ** -24,) if args.reinject else (),
                                            checkpoints=[o * k / 10 for k in range(1, 10)])}))
        return
    cells = params.nrad * params.nsec
    prof_steps = min(args.steps, 60)
    res = {}
    for prec, name in ((_capi.FP64, "fp64"), (_capi.FP32, "fp32")):
        print(f"bench: {args.workload} {name}", file=sys.stderr, flush=True)
        el, kt = run_gpu(params, planets, prec, args.steps, args.warmup, prof_steps, dust)
        res[name] = dict(elapsed=el, kt=kt)
    w32, w64 = 4, 8
    energy, sg = bool(params.energy_equation), bool(params.self_gravity)

    def summary(name, w):
        el, kt = res[name]["elapsed"], res[name]["kt"]
        value = args.steps * cells / el
        ms = el / args.steps * 1e3
        # the roofline kernel: the largest share of step time among kernels
        # that move field data (the single-block control kernel moves none)
        nbytes = lambda k: kernel_bytes_per_launch(k, w, energy, sg, cells, npart)
        dom = max(((k, v) for k, v in kt.items() if nbytes(k) > 0), key=lambda kv: kv[1][0])
        kname, (tot_ms, nl) = dom
        avg_s = tot_ms / nl * 1e-3
        kbytes = nbytes(kname)
        sbytes = step_bytes_per_cell(w, energy, sg) * cells
        prof_total = sum(v[0] for v in kt.values()) / prof_steps
        kernels = {k: {"avg_us": round(v[0] / v[1] * 1e3, 3), "launches": int(v[1]),
                       "share": round(v[0] / prof_steps / prof_total, 4)}
                   for k, v in sorted(kt.items(), key=lambda kv: -kv[1][0])}
        return value, ms, kname, avg_s, kbytes, sbytes, kernels

    v32, ms32, k32, a32, kb32, sb32, kern32 = summary("fp32", w32)
    v64, ms64, k64, a64, kb64, sb64, kern64 = summary("fp64", w64)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}_{params.nrad}x{params.nsec}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        entry = tj.get("fp32", {}).get(k32)
        # HBM bytes per launch of the roofline kernel (FETCH_SIZE + WRITE_SIZE,
        # scaled by a plain copy of known size)
        traffic = entry["bytes"] if entry else None
    # the roofline of the dominant kernel: the sub-step's algorithmic bytes
    # per cell times the cells of one launch over its average duration; and
    # the same with the kernel's own compulsory bytes
    ach = sb32 / a32 / 1e9
    ach_own = kb32 / a32 / 1e9
    line = {
        "metric": "cell-updates/s per hydro step (% HBM roofline), FP32 vs FP64 at <=1e-5 error",
        "value": v32,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "storage_dtype": "f32",
        "data": "synthetic (seeded AddNoise initial disc)",
        "config": {"workload": f"{args.workload} {params.nrad}x{params.nsec}",
                   "nrad": params.nrad, "nsec": params.nsec,
                   "energy_equation": energy, "self_gravity": sg,
                   "alpha_viscosity": params.alpha_viscosity, "planets": len(planets["dist"]),
                   "particles": npart},
        "roofline": {"bound": "hbm", "kernel": k32, "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": sb32, "avg_launch_us": a32 * 1e6,
                     "kernel_own_bytes_per_launch": kb32, "kernel_own_frac": ach_own / HBM_PEAK_GBS},
        "roofline_step": {"bytes_per_cell": step_bytes_per_cell(w32, energy, sg),
                          "achieved": sb32 / (ms32 * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": sb32 / (ms32 * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "fp64": {"value": v64, "ms_per_step": ms64, "dominant_kernel": k64,
                 "roofline_step_frac": sb64 / (ms64 * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "fp32_over_fp64": v32 / v64,
        "kernels_fp32": kern32,
        "kernels_fp64": kern64,
        "cpu_baseline": None,
    }
    if args.target_steps > 0 and args.workload == "energy_sg":
        # the 1024x3072 FP32 gas+SG disc (with a planet), 40 B per cell
        from dustyfargoadsg_amd.workloads import gas_sg_large
        print("bench: gas_sg_large 1024x3072", file=sys.stderr, flush=True)
        tp, tpl = gas_sg_large()
        tcells = tp.nrad * tp.nsec
        el32, _ = run_gpu(tp, tpl, _capi.FP32, args.target_steps, 20, 1)
        el64, _ = run_gpu(tp, tpl, _capi.FP64, args.target_steps, 20, 1)
        tb = step_bytes_per_cell(4, False, True)
        ach_t = tb * tcells * args.target_steps / el32 / 1e9
        line["target_1024x3072"] = {
            "workload": "gas_sg_large 1024x3072 (isothermal gas + SG + planet)", "steps": args.target_steps,
            "value": args.target_steps * tcells / el32, "ms_per_step": el32 / args.target_steps * 1e3,
            "fp64_value": args.target_steps * tcells / el64, "fp32_over_fp64": el64 / el32,
            "roofline_step": {"bytes_per_cell": tb, "achieved": ach_t, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": ach_t / HBM_PEAK_GBS}}
    if args.batched > 0 and args.workload == "energy_sg":
        print(f"bench: batched sweep of {args.batched} discs", file=sys.stderr, flush=True)
        line["batched"] = batched_sweep(args.batched, 4, args.batched_calls)
    if args.error_orbits > 0 and dust is None:
        line["fp32_error_vs_fp64"] = precision_error(params, planets, args.error_orbits, control=False)
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cv, n, el = cpu_baseline(params, planets, args.cpu_seconds, dust)
        line["cpu_baseline"] = {"value": cv, "unit": "cell-updates/s", "cores": 1, "kind": "port",
                                "sample": f"{n} sub-steps of the same {params.nrad}x{params.nsec} "
                                          f"disc by the FP64 C oracle, {el:.1f} s, 1 thread"}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
