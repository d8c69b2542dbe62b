#!/usr/bin/env python3
"""Hybrid UVLM-VPM steps/s and G pair-interactions/s on one MI355X.

Workload: the heaving-wing hybrid case (rectangular NACA 0012 wing, chord
1 m, half span 5 m, 10 x 12 panels, V = 10 m/s, h = 0.1 c sin(2 k V t / c) at
k = 0.5, dt = period/50, N_keep = 10 ring rows, older rows converted to
Winckelmans particles with sigma = 0.5 V dt, free-wake self-induction +
rVPM stretching + Pedrizzetti relaxation).  The case is run for 3 cycles, the
particle wake is then extended periodically downstream to the target count
(synthetic far wake of the same shape), and K full hybrid steps are timed
with HIP events on the library's stream.

    python bench.py [--steps K] [--warmup W] [--particles N] [--gpus 1]

Prints one JSON line (rank 0): value = FP64 steps/s at --particles particles.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

# Algorithmic flops per pair (Winckelmans velocity 31, fused velocity +
# gradient 66, Gaussian-erf fused 72, Ptera-form segment 51).
FLOPS = {"particle_vel": 31, "particle_stretch": 66, "particle_curl": 72,
         "particle_induce": 31, "seg_pairs": 51}
PEAK_FP64 = 78.6e12
PEAK_FP32 = 157.3e12


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_case(hybrid, steps_total, precision, max_particles):
    return hybrid.heaving_wing_case(num_steps=steps_total, precision=precision,
                                    max_particles=max_particles)


def extend_wake(state, target, period_shift, period_age, block):
    """Appends downstream copies of the oldest `block` particles (one heave
    period of wake) shifted by k * V * T until `target` particles exist."""
    pos, gam, sig, age = state["pos"], state["gamma"], state["sigma"], state["age"]
    n0 = pos.shape[0]
    if target <= n0:
        return {k: v[:target].copy() for k, v in state.items()}
    base = slice(0, min(block, n0))
    need = target - n0
    reps = []
    k = 1
    while need > 0:
        take = min(need, base.stop)
        p = pos[base][:take].copy()
        p[:, 0] += k * period_shift
        reps.append((p, gam[base][:take], sig[base][:take], age[base][:take] + k * period_age))
        need -= take
        k += 1
    return {
        "pos": np.concatenate([pos] + [r[0] for r in reps]),
        "gamma": np.concatenate([gam] + [r[1] for r in reps]),
        "sigma": np.concatenate([sig] + [r[2] for r in reps]),
        "age": np.concatenate([age] + [r[3] for r in reps]),
    }


def measure(hybrid, lib, particles, steps, warmup, precision, prefill=150, timing_steps=2,
            seed_state=None):
    """Returns a dict with the timed-region and per-kernel measurements."""
    total = prefill + warmup + steps + timing_steps + 2
    cap = particles + 64 * total
    case = build_case(hybrid, total, precision, cap)
    run = hybrid.HybridRun(case)
    ctx = run._ctx
    run.step(prefill)
    st = run.particles.state_dict() if seed_state is None else seed_state
    T = case.meta["period"]
    per_period = 50 * 2 * 6 * 4          # rows per period x rings per row x 4
    st = extend_wake(st, particles, case.meta["V"] * T, T, per_period)
    run.particles.load_state_dict(st)
    run.step(warmup)
    ctx.synchronize()
    seg0, par0 = run.pair_counts()
    np0 = run.particles.np
    t0 = time.perf_counter()
    ctx.event_record(0)
    run.step(steps)
    ctx.event_record(1)
    ms = ctx.event_elapsed_ms(0, 1)
    wall = time.perf_counter() - t0
    seg1, par1 = run.pair_counts()
    np1 = run.particles.np
    cl, cdi, _ = run.read()
    s0 = prefill + warmup
    cl_t = cl[s0:s0 + steps].copy()
    # per-kernel timing pass (events around every launch of each family)
    ctx.set_timing(True)
    fam = {}
    for _ in range(timing_steps):
        run.step(1)
        n = run.particles.np
        for f in ("particle_vel", "particle_stretch", "particle_curl", "particle_induce",
                  "seg_pairs", "aic", "dense_solve"):
            k, t = ctx.kernel_stats(f)
            e = fam.setdefault(f, {"launches": 0, "ms": 0.0, "pairs": 0.0})
            e["launches"] += k - e.get("_k", 0)
            e["ms"] += t - e.get("_t", 0.0)
            e["_k"], e["_t"] = k, t
            if f.startswith("particle_") and f != "particle_induce":
                e["pairs"] += (k - e.get("_kp", 0)) * float(n) * float(n)
                e["_kp"] = k
    ctx.set_timing(False)
    for e in fam.values():
        for key in ("_k", "_t", "_kp"):
            e.pop(key, None)
    run.close()
    return {"ms_total": ms, "wall_s": wall, "steps": steps, "np_start": np0, "np_end": np1,
            "seg_pairs": seg1 - seg0, "particle_pairs": par1 - par0, "cl": cl_t,
            "families": fam, "seed_state": st}


def cpu_baseline(particles, threads):
    """Times the oracle's particle sums (the step's dominant work) on a bounded
    target subset with all sources, and scales linearly to a whole step."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    rng = np.random.default_rng(0)
    n = particles
    pos = rng.uniform([0, -5, -1], [200, 5, 1], (n, 3))
    al = rng.normal(0, 0.01, (n, 3))
    sg = np.full(n, 0.0628)
    sub = min(n, 2048)
    tgt = pos[:sub]
    oracle.particle_vel(tgt[:16], pos, al, sg, "winckelmans", threads)   # warm-up
    t0 = time.perf_counter()
    oracle.particle_vel(tgt, pos, al, sg, "winckelmans", threads)
    t_vel = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.particle_jac(tgt, pos, al, sg, "winckelmans", threads)
    t_jw = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.particle_jac(tgt, pos, al, sg, "gaussianerf", threads)
    t_jg = time.perf_counter() - t0
    scale = n / sub
    t_step = scale * (3 * t_vel + 3 * t_jw + t_jg)
    return {"value": 1.0 / t_step, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": (f"oracle particle sums (3 Winckelmans velocity + 3 Winckelmans Jacobian "
                       f"+ 1 Gaussian Jacobian per step) for {sub} of {n} targets x {n} sources, "
                       f"scaled x{scale:.1f}; ring/UVLM part (<0.1%) omitted"),
            "sample_seconds": t_vel + t_jw + t_jg}


def roofline(fam, precision):
    """Dominant particle kernel: algorithmic flops per launch / average launch time."""
    cands = {k: v for k, v in fam.items() if k in ("particle_vel", "particle_stretch",
                                                    "particle_curl") and v["launches"]}
    name = max(cands, key=lambda k: cands[k]["ms"])
    e = cands[name]
    avg_s = e["ms"] / e["launches"] / 1e3
    flops = FLOPS[name] * e["pairs"] / e["launches"]
    achieved = flops / avg_s
    peak = PEAK_FP64
    return name, {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12,
                  "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                  "kernel": f"k_pairs<PartOp> ({name})", "avg_launch_ms": avg_s * 1e3,
                  "flops_per_launch": flops,
                  "note": ("FP64 VALU pair sum; MI355X FP64 vector peak = FP64 matrix peak = "
                           "78.6 TFLOP/s" if precision == "fp64" else
                           "FP32 pair terms with FP64 accumulation, priced at the FP64 peak")}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--particles", type=int, default=20000)
    ap.add_argument("--big", type=int, default=100000,
                    help="second particle count measured for FP64 vs mixed (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": a.gpus}))
        return 2
    from fluxvortex_amd import _lib, hybrid
    lib = _lib.load_library()
    log(f"[bench] heaving wing, {a.particles} particles, FP64, warmup {a.warmup}, steps {a.steps}")
    m64 = measure(hybrid, lib, a.particles, a.steps, a.warmup, "fp64")
    sps = a.steps / (m64["ms_total"] / 1e3)
    pairs_total = m64["seg_pairs"] + m64["particle_pairs"]
    name, roof = roofline(m64["families"], "fp64")
    out = {
        "metric": "hybrid UVLM-VPM steps/s & G pair-interactions/s: FP64 vs mixed+learned",
        "value": sps, "unit": "steps/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": m64["ms_total"] / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "heaving-wing hybrid UVLM-VPM step (120 panels, N_keep=10, "
                               "Winckelmans free wake + stretching + relaxation)",
                   "particles": [m64["np_start"], m64["np_end"]], "panels": 120,
                   "precision": "fp64", "parallelism": "single GPU"},
        "gpairs_per_s": pairs_total / (m64["ms_total"] / 1e3) / 1e9,
        "segment_gpairs_per_s": m64["seg_pairs"] / (m64["ms_total"] / 1e3) / 1e9,
        "particle_gpairs_per_s": m64["particle_pairs"] / (m64["ms_total"] / 1e3) / 1e9,
        "roofline": roof,
        "kernel_families_ms": {k: round(v["ms"] / max(v["launches"], 1), 4)
                               for k, v in m64["families"].items()},
    }
    if not a.no_mixed:
        log("[bench] same workload, mixed precision")
        mmx = measure(hybrid, lib, a.particles, a.steps, a.warmup, "mixed",
                      seed_state=m64["seed_state"])
        scl = float(np.max(np.abs(m64["cl"])))
        out["mixed"] = {"steps_per_s": a.steps / (mmx["ms_total"] / 1e3),
                        "speedup_vs_fp64": m64["ms_total"] / mmx["ms_total"],
                        "cl_max_rel_err_vs_fp64": float(np.max(np.abs(mmx["cl"] - m64["cl"])))
                        / scl}
    if a.big:
        log(f"[bench] {a.big} particles, FP64")
        b64 = measure(hybrid, lib, a.big, 2, 1, "fp64", timing_steps=1)
        big = {"particles": b64["np_start"], "fp64_steps_per_s": 2 / (b64["ms_total"] / 1e3),
               "fp64_gpairs_per_s": (b64["seg_pairs"] + b64["particle_pairs"])
               / (b64["ms_total"] / 1e3) / 1e9}
        if not a.no_mixed:
            log(f"[bench] {a.big} particles, mixed")
            bmx = measure(hybrid, lib, a.big, 2, 1, "mixed", timing_steps=1,
                          seed_state=b64["seed_state"])
            big["mixed_steps_per_s"] = 2 / (bmx["ms_total"] / 1e3)
            big["mixed_speedup_vs_fp64"] = b64["ms_total"] / bmx["ms_total"]
            big["mixed_cl_max_rel_err_vs_fp64"] = float(
                np.max(np.abs(bmx["cl"] - b64["cl"])) / np.max(np.abs(b64["cl"])))
        out["heaving_100k"] = big
    if not a.no_cpu_baseline:
        threads = os.cpu_count() or 1
        log(f"[bench] CPU baseline (oracle, {threads} threads)")
        cb = cpu_baseline(a.particles, threads)
        cb["cpu_model"] = platform.processor() or platform.machine()
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
