#!/usr/bin/env python3
"""Hybrid UVLM-VPM steps/s and G pair-interactions/s on one MI355X.

Workload: the heaving-wing hybrid case (rectangular NACA 0012 wing, chord
1 m, half span 5 m, 10 x 12 panels, V = 10 m/s, h = 0.1 c sin(2 k V t / c) at
k = 0.5, dt = period/50, N_keep = 10 ring rows, older rows converted to
Winckelmans particles with sigma = 0.5 V dt, free-wake self-induction +
rVPM stretching + Pedrizzetti relaxation).  The case is run for 3 cycles with
the particle wake free but unstretched (with stretching this configuration
diverges within ~25 steps in the reference algorithm itself), the particle
wake is then extended periodically downstream to the target count (synthetic
far wake of the same shape), and K full hybrid steps (stretching on) are timed
with HIP events on the library's stream.  Accuracy comparisons (mixed,
mixed + learned closure) start from bitwise-identical states and use the
unstretched wake, whose history is well conditioned.

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
BIG_STEPS = 10   # timed steps of the 100k block: one closure check period
PEAK_FP32 = 157.3e12
# HBM bytes per launch of the stretching pair sum at the default 20k workload
# (20 096 particles at the timed steps, 64 lanes x 2 targets, 27 source slices),
# from two rocprofv3 counter passes
# (FETCH_SIZE x 2.00, WRITE_SIZE x 1.00, scaled by a plain 512 MiB copy):
# profiles/k_pairs_stretch_heaving_wing_20k_pmc.md
TRAFFIC_STRETCH = {"particles": 20096, "bytes": (4187.5 * 2.0 + 12717.0) * 1024}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def extend_wake(state, target, period_shift, period_age, block):
    """Appends downstream copies of the oldest `block` particles (one heave
    period of wake) shifted by k * V * T until `target` particles exist."""
    pos, gam, sig, age = state["pos"], state["gamma"], state["sigma"], state["age"]
    n0 = pos.shape[0]
    if target <= n0:
        return {k: v[:target].copy() for k, v in state.items()}
    base = min(block, n0)
    need = target - n0
    reps = []
    k = 1
    while need > 0:
        take = min(need, base)
        p = pos[:take].copy()
        p[:, 0] += k * period_shift
        reps.append((p, gam[:take], sig[:take], age[:take] + k * period_age))
        need -= take
        k += 1
    return {
        "pos": np.concatenate([pos] + [r[0] for r in reps]),
        "gamma": np.concatenate([gam] + [r[1] for r in reps]),
        "sigma": np.concatenate([sig] + [r[2] for r in reps]),
        "age": np.concatenate([age] + [r[3] for r in reps]),
    }


PREFILL = 150


def prepare(hybrid, particles, steps_after):
    """A heaving-wing run advanced 3 cycles with the particle wake free but
    unstretched (the reference HybridSolver's own advect, which stays physical),
    its particle wake then extended to `particles`.  Deterministic: two calls
    give bitwise-identical states."""
    total = PREFILL + steps_after + 2
    case = hybrid.heaving_wing_case(num_steps=total, precision="fp64",
                                    max_particles=particles + 64 * total)
    run = hybrid.HybridRun(case)
    run.set_flags(True, False)
    run.step(PREFILL)
    T = case.meta["period"]
    st = extend_wake(run.particles.state_dict(), particles, case.meta["V"] * T, T,
                     50 * 2 * 6 * 4)
    run.particles.load_state_dict(st)
    return run, case


def timed(run, steps):
    ctx = run._ctx
    ctx.synchronize()
    seg0, par0 = run.pair_counts()
    np0 = run.particles.np
    s0 = run.current_step
    ctx.event_record(0)
    run.step(steps)
    ctx.event_record(1)
    ms = ctx.event_elapsed_ms(0, 1)
    seg1, par1 = run.pair_counts()
    cl, cdi, _ = run.read()
    return {"ms": ms, "steps": steps, "np": [np0, run.particles.np], "seg_pairs": seg1 - seg0,
            "particle_pairs": par1 - par0, "cl": cl[s0:s0 + steps].copy()}


def kernel_families(run, steps=2):
    """Per-launch HIP-event timing of each kernel family over `steps` steps."""
    ctx = run._ctx
    ctx.set_timing(True)
    fam = {}
    for _ in range(steps):
        marks = {f: ctx.kernel_stats(f) for f in FAMILIES}
        run.step(1)
        n = run.particles.np
        for f in FAMILIES:
            k, t = ctx.kernel_stats(f)
            e = fam.setdefault(f, {"launches": 0, "ms": 0.0, "pairs": 0.0})
            e["launches"] += k - marks[f][0]
            e["ms"] += t - marks[f][1]
            if f in SELF_FAMILIES:
                e["pairs"] += (k - marks[f][0]) * float(n) * float(n)
    ctx.set_timing(False)
    return fam


FAMILIES = ("particle_vel", "particle_stretch", "particle_curl", "particle_induce", "seg_pairs",
            "aic", "dense_solve", "closure_near", "closure_far", "closure_train",
            "closure_check", "mlp_forward", "mlp_train")
SELF_FAMILIES = ("particle_vel", "particle_stretch", "particle_curl")


def run_config(hybrid, particles, steps, warmup, precision, stretch=True, closure=None):
    run, case = prepare(hybrid, particles, warmup + steps + 3)
    if precision != "fp64":
        run.set_precision(precision)
    net = None
    if closure is not None:
        from fluxvortex_amd.mlp import ClosureMLP, closure_params
        net = ClosureMLP(closure_params(seed=0))
        clo = dict(closure)
        kl = clo.pop("kappa_local", 0.0)
        run.particles.enable_closure(net, **clo)
        if kl:
            run.particles.set_closure_local(kl)
        run.particles.closure_train(rounds=800)
    run.set_flags(True, stretch)
    run.step(warmup)
    out = timed(run, steps)
    out["families"] = kernel_families(run, 1 if particles > 50000 else 2)
    if closure is not None:
        out["closure"] = run.particles.closure_stats()
    run.close()
    return out


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
    sub = n
    tgt = pos[:sub]
    oracle.particle_vel(tgt[:16], pos, al, sg, "winckelmans", threads)   # warm-up

    def best(fn, *args, reps=3):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn(*args)
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), sum(ts)

    t_vel, s1 = best(oracle.particle_vel, tgt, pos, al, sg, "winckelmans", threads)
    t_jw, s2 = best(oracle.particle_jac, tgt, pos, al, sg, "winckelmans", threads)
    t_jg, s3 = best(oracle.particle_jac, tgt, pos, al, sg, "gaussianerf", threads)
    scale = n / sub
    t_step = scale * (3 * t_vel + 3 * t_jw + t_jg)
    return {"value": 1.0 / t_step, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": (f"oracle particle sums of one step (3 Winckelmans velocity + 3 Winckelmans "
                       f"Jacobian + 1 Gaussian Jacobian) for all {n} targets x {n} sources, each "
                       f"sum the median of 3 runs; ring/UVLM part (<0.1%) omitted"),
            "sample_seconds": s1 + s2 + s3}


def roofline(fam, precision, particles=None):
    """Dominant particle kernel: algorithmic flops per launch / average launch time."""
    cands = {k: v for k, v in fam.items() if k in SELF_FAMILIES and v["launches"]}
    name = max(cands, key=lambda k: cands[k]["ms"])
    e = cands[name]
    avg_s = e["ms"] / e["launches"] / 1e3
    flops = FLOPS[name] * e["pairs"] / e["launches"]
    achieved = flops / avg_s
    peak = PEAK_FP64
    traffic = None
    if name == "particle_stretch" and precision == "fp64" and particles == TRAFFIC_STRETCH["particles"]:
        traffic = TRAFFIC_STRETCH["bytes"]
    return name, {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12,
                  "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                  "kernel": f"k_pairs<PartOp> ({name})", "avg_launch_ms": avg_s * 1e3,
                  "flops_per_launch": flops,
                  "note": ("FP64 VALU pair sum; MI355X FP64 vector peak = FP64 matrix peak = "
                           "78.6 TFLOP/s" if precision == "fp64" else
                           "FP32 pair terms with FP64 accumulation, priced at the FP64 peak")}


def flutter_block(n_speeds=64, cpu_sample=4):
    """The batched Goland flutter sweep: n_speeds airspeeds uniform over
    80-180 m/s, 100 chords of wake, N_keep 10 (uneven step and particle
    counts), every instance stepped together (one induction + one step launch
    per time step, replayed from a hipGraph).  FP64 and mixed; then the
    reference benchmark's own wake (every row kept as rings, held in device
    memory, one more launch per step) in FP64 against its published flutter
    speed.  The oracle is timed on cpu_sample speeds of the N_keep 10 sweep
    and scaled to the sweep."""
    from fluxvortex_amd import aeroelastic as ae
    speeds = np.linspace(80.0, 180.0, n_speeds)
    res = {"speeds": [80.0, 180.0, n_speeds], "n_keep": 10, "num_chords": 100}
    for prec, n_keep in (("fp64", 10), ("mixed", 10), ("ring_wake_fp64", 100000)):
        insts = [ae.goland_instance(float(V), precision=prec.replace("ring_wake_", ""),
                                    n_keep=n_keep) for V in speeds]
        sw = ae.Sweep(insts, use_graph=True)
        sw.step(1)
        sw.read(0)
        t0 = time.perf_counter()
        sw.step()
        sw.read(0)
        sec = time.perf_counter() - t0
        sig = []
        inst_steps = 0
        for i, it in enumerate(insts):
            r = sw.read(i)
            if r["status"] != 0:
                raise RuntimeError(f"flutter instance {i} failed ({r['status']})")
            sig.append(ae.envelope_growth(r["tip_w"][~np.isnan(r["tip_w"])], it.dt))
            inst_steps += it.num_steps - 1
        sw.close()
        res[prec] = {"seconds": sec, "time_steps": sw.max_steps - 1,
                     "instance_steps_per_s": inst_steps / sec,
                     "flutter_speed": ae.flutter_speed(speeds, sig),
                     "sigma_w_80": sig[0], "sigma_w_180": sig[-1]}
    # benchmark_goland.py's own eight speeds and published result (README.md:121-132)
    pub_V = [80.0, 100.0, 120.0, 130.0, 140.0, 144.0, 160.0, 180.0]
    pub = ae.flutter_sweep(pub_V, use_graph=True, n_keep=100000)
    res["ring_wake_published_speeds"] = {
        "speeds": pub_V, "sigma_w": [r["sigma_w"] for r in pub["runs"]],
        "flutter_speed": pub["flutter_speed"], "published_flutter_speed": 140.2,
        "published_sigma_w": [-0.628, -0.899, -0.249, -0.101, -0.037, 0.564, 0.794, 1.413]}
    f64, mx = res["fp64"]["flutter_speed"], res["mixed"]["flutter_speed"]
    if f64 and mx:
        res["mixed_flutter_speed_rel_err_vs_fp64"] = abs(mx - f64) / f64
    try:
        sys.path.insert(0, str(ROOT / "oracle"))
        import oracle
        pick = np.linspace(0, n_speeds - 1, cpu_sample).round().astype(int)
        t0 = time.perf_counter()
        for k in pick:
            oracle.aeroelastic_run(ae.goland_instance(float(speeds[k])), threads=1)
        per = (time.perf_counter() - t0) / cpu_sample
        res["cpu_oracle"] = {"seconds_est": per * n_speeds, "threads": 1,
                             "sample": f"{cpu_sample} of the {n_speeds} speeds, scaled"}
    except Exception as e:   # the oracle is optional on a user's machine
        res["cpu_oracle"] = {"error": str(e)}
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--particles", type=int, default=20000)
    ap.add_argument("--big", type=int, default=100000,
                    help="second particle count: FP64 vs mixed vs mixed+learned (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-flutter", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": a.gpus}))
        return 2
    from fluxvortex_amd import _lib, hybrid
    _lib.load_library()
    # tile shapes of the pair sums (and the closure's near field) timed at
    # first use, for every configuration below alike
    _lib.default_context().set_autotune(not a.no_autotune)
    log(f"[bench] heaving wing, {a.particles} particles, FP64, warmup {a.warmup}, steps {a.steps}")
    m64 = run_config(hybrid, a.particles, a.steps, a.warmup, "fp64")
    sec = m64["ms"] / 1e3
    pairs_total = m64["seg_pairs"] + m64["particle_pairs"]
    np_timed = m64["np"][0] if isinstance(m64["np"], (list, tuple)) else m64["np"]
    name, roof = roofline(m64["families"], "fp64", np_timed)
    out = {
        "metric": "hybrid UVLM-VPM steps/s & G pair-interactions/s: FP64 vs mixed+learned",
        "value": a.steps / sec, "unit": "steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": m64["ms"] / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "heaving-wing hybrid UVLM-VPM step: 120 panels, N_keep=10, "
                               "Winckelmans sigma=0.5 V dt, free wake + rVPM stretching + "
                               "relaxation, RK3",
                   "particles": m64["np"], "panels": 120, "precision": "fp64",
                   "parallelism": "single GPU"},
        "gpairs_per_s": pairs_total / sec / 1e9,
        "segment_gpairs_per_s": m64["seg_pairs"] / sec / 1e9,
        "particle_gpairs_per_s": m64["particle_pairs"] / sec / 1e9,
        "roofline": roof,
        "kernel_family_ms_per_launch": {k: round(v["ms"] / v["launches"], 4)
                                        for k, v in m64["families"].items() if v["launches"]},
    }
    if not a.no_mixed:
        log("[bench] same workload, mixed precision")
        mmx = run_config(hybrid, a.particles, a.steps, a.warmup, "mixed")
        out["mixed"] = {"steps_per_s": a.steps / (mmx["ms"] / 1e3),
                        "speedup_vs_fp64": m64["ms"] / mmx["ms"]}
        log("[bench] accuracy from identical states (unstretched free wake)")
        r64 = run_config(hybrid, a.particles, a.steps, a.warmup, "fp64", stretch=False)
        rmx = run_config(hybrid, a.particles, a.steps, a.warmup, "mixed", stretch=False)
        scl = float(np.max(np.abs(r64["cl"])))
        out["mixed"]["cl_max_rel_err_vs_fp64"] = float(np.max(np.abs(rmx["cl"] - r64["cl"]))) / scl
    if a.big:
        # BIG_STEPS timed steps after one warm-up: the closure checks itself
        # against the exact direct sum every 10 steps, so the window holds
        # exactly one check and its cost enters the mean step
        big = {"timed_steps": BIG_STEPS}
        log(f"[bench] {a.big} particles, FP64")
        b64 = run_config(hybrid, a.big, BIG_STEPS, 1, "fp64")
        big["particles"] = b64["np"]
        big["fp64_steps_per_s"] = BIG_STEPS / (b64["ms"] / 1e3)
        big["fp64_gpairs_per_s"] = (b64["seg_pairs"] + b64["particle_pairs"]) / (b64["ms"] / 1e3) / 1e9
        if not a.no_mixed:
            log(f"[bench] {a.big} particles, mixed")
            bmx = run_config(hybrid, a.big, BIG_STEPS, 1, "mixed")
            big["mixed_steps_per_s"] = BIG_STEPS / (bmx["ms"] / 1e3)
            # the cheapest (cell edge, cutoff) whose closure passes every one of
            # its own error checks in the run (a rejected check sends the steps
            # back to the direct sums; such a candidate is not kept); the near
            # field reaches cell * kappa, the local expansions begin at twice that
            tried = []
            for cell, kappa in ((2.5, 2.0), (2.0, 2.5), (2.0, 3.0), (2.0, 4.0)):
                log(f"[bench] {a.big} particles, mixed + learned closure, cell {cell} kappa {kappa}")
                clo = dict(cell=cell, kappa=kappa, check_every=10, tol=1e-3, train_samples=4096,
                           train_iters=2, lr=2e-3, kappa_local=2.0 * kappa)
                bml = run_config(hybrid, a.big, BIG_STEPS, 1, "mixed", closure=clo)
                st = bml["closure"]
                tried.append({"cell": cell, "kappa": kappa, "checks": st["checks"],
                              "rejections": st["rejections"], "last_err": st["last_err"],
                              "steps_per_s": BIG_STEPS / (bml["ms"] / 1e3)})
                if st["active"] and st["rejections"] == 0:
                    break
            big["closure_candidates"] = tried
            big["mixed_learned_steps_per_s"] = BIG_STEPS / (bml["ms"] / 1e3)
            big["mixed_learned_speedup_vs_fp64"] = b64["ms"] / bml["ms"]
            big["closure"] = bml["closure"]
            big["closure_params"] = clo
            log("[bench] closure accuracy from identical states (unstretched free wake)")
            r64 = run_config(hybrid, a.big, 3, 1, "fp64", stretch=False)
            rml = run_config(hybrid, a.big, 3, 1, "mixed", stretch=False, closure=clo)
            big["mixed_learned_cl_max_rel_err_vs_fp64"] = float(
                np.max(np.abs(rml["cl"] - r64["cl"])) / np.max(np.abs(r64["cl"])))
        out["heaving_100k"] = big
    if not a.no_flutter:
        log("[bench] batched Goland flutter sweep, 64 speeds")
        out["flutter_sweep"] = flutter_block()
    out["autotune"] = [dict(zip(("family", "log2_targets", "log2_sources", "lanes",
                                 "targets_per_lane", "workgroups_per_cu", "best_ms"), e))
                       for e in _lib.default_context().autotune_report()
                       if e[1] >= 14]
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
