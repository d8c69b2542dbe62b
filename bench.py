#!/usr/bin/env python3
"""Hybrid UVLM-VPM steps/s and G pair-interactions/s on one MI355X.

Workload (`value`): the heaving-wing hybrid case (rectangular NACA 0012 wing,
chord 1 m, half span 5 m, 10 x 12 panels, V = 10 m/s, h = 0.1 c sin(2 k V t /
c) at k = 0.5, dt = period/50, N_keep = 10 ring rows, older rows converted to
Winckelmans particles with sigma = 0.5 V dt, free-wake self-induction + rVPM
stretching + Pedrizzetti relaxation under the reference CPU path's pair rule).
The case is run for 3 cycles with the particle wake free but unstretched (the
reference HybridSolver's own advect), the particle wake is then extended
periodically downstream to the target count (synthetic far wake of the same
shape), and K full hybrid steps (stretching on) are timed with HIP events on
the library's stream; every input is resident in HBM.

Beside it, one JSON object each: mixed precision and its accuracy against the
FP64 step's own sensitivity; the 100k-particle block (FP64, mixed, mixed +
learned closure, the closure without its network); per-kernel rooflines; the
reference's published collapsed-ring routine sizes through the patch surface;
the static wing (C1), the seeded raw-kernel sweep (C3), the flapping airplane
with the closure (C5); the batched Goland flutter sweep (C4); the CPU oracle.

    python bench.py [--steps K] [--warmup W] [--particles N] [--gpus 1]

Prints one JSON line: value = FP64 steps/s at --particles particles.
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
PEAK_FP64 = 78.6e12          # MI355X FP64 vector (= FP64 matrix) peak
PEAK_FP32_MFMA = 157.3e12
PEAK_BF16 = 2.5e15           # dense BF16 MFMA
PEAK_HBM = 8.0e12
MLP_FLOPS_PER_SAMPLE = 2 * (32 * 128 + 2 * 128 * 128 + 128 * 16)   # 77 824
BIG_STEPS = 10   # timed steps of the 100k block: one closure check period
# heave cycles of free, unstretched wake before a timed window: nine reach
# ~21k particles by themselves (the 20k line); larger states copy periods
NATURAL_CYCLES = 9
# HBM fetch of the stretching pair sum per target particle: FETCH_SIZE x 2.00
# (the counter's scale on a plain copy) over the launches of the profiled
# 20k bench, divided by their particle count (profiles/
# k_pairs_stretch_heaving_wing_20k_pmc.md); the writes are the split's
# partial sums, split x N x 3 x 8 B, which the WRITE_SIZE pass confirms.
FETCH_BYTES_PER_PARTICLE = 4187.5 * 1024 * 2.0 / 20096


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def extend_wake(state, target, period_shift, period_age, block, period=0):
    """Appends downstream copies of one heave period of wake (the `block`
    particles of the `period`-th oldest period) shifted by (k + period) V T,
    k = 1, 2, ..., until `target` particles exist."""
    pos, gam, sig, age = state["pos"], state["gamma"], state["sigma"], state["age"]
    n0 = pos.shape[0]
    if target <= n0:
        return {k: v[:target].copy() for k, v in state.items()}
    p0 = min(period * block, max(n0 - block, 0))
    base = min(block, n0 - p0)
    src = slice(p0, p0 + base)
    need = target - n0
    reps = []
    k = 1
    while need > 0:
        take = min(need, base)
        p = pos[src][:take].copy()
        p[:, 0] += (k + period) * period_shift
        reps.append((p, gam[src][:take], sig[src][:take], age[src][:take] + (k + period) * period_age))
        need -= take
        k += 1
    return {
        "pos": np.concatenate([pos] + [r[0] for r in reps]),
        "gamma": np.concatenate([gam] + [r[1] for r in reps]),
        "sigma": np.concatenate([sig] + [r[2] for r in reps]),
        "age": np.concatenate([age] + [r[3] for r in reps]),
    }


def prepare(hybrid, particles, steps_after, perturb=0.0):
    """A heaving-wing run advanced with the particle wake free but unstretched
    (the reference HybridSolver's advect) for as many steps as it takes the
    wake to reach `particles` by itself (no particles for the first N_keep +
    1 = 11 steps, then 48 a step: 428 steps = 8.6 heave cycles for 20 000; at
    most NATURAL_CYCLES cycles, at least 3), its particle wake then extended
    to `particles`, if it is still short, by downstream copies of its
    second-oldest heave period; alpha scaled by 1 + perturb.
    Deterministic: two calls give bitwise-identical states."""
    prefill = min(50 * NATURAL_CYCLES, max(150, 11 + -(-particles // 48)))
    total = prefill + steps_after + 2
    case = hybrid.heaving_wing_case(num_steps=total, precision="fp64",
                                    max_particles=particles + 64 * total)
    run = hybrid.HybridRun(case)
    run.set_flags(True, False)
    run.step(prefill)
    T = case.meta["period"]
    st = run.particles.state_dict()
    if st["pos"].shape[0] < particles:
        # copies of the second-oldest period (the oldest holds the start-up
        # transient), placed downstream of the oldest
        st = extend_wake(st, particles, case.meta["V"] * T, T, 50 * 2 * 6 * 4, period=1)
    if perturb:
        st["gamma"] = st["gamma"] * (1.0 + perturb)
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


FAMILIES = ("particle_vel", "particle_stretch", "particle_curl", "particle_induce", "seg_pairs",
            "aic", "dense_solve", "convert", "closure_near", "closure_far", "closure_train",
            "closure_residual", "mlp_forward", "mlp_train", "mlp_adam")
SELF_FAMILIES = ("particle_vel", "particle_stretch", "particle_curl")


def kernel_families(run, steps=2):
    """Per-launch HIP-event timing of each kernel family over `steps` steps
    (events on the stream each kernel is launched on)."""
    ctx = run._ctx
    ctx.set_timing(True)
    fam = {}
    clo = run.particles.has_closure()
    for _ in range(steps):
        marks = {f: ctx.kernel_stats(f) for f in FAMILIES}
        near0 = run.particles.closure_stats()["near_pairs"] if clo else 0.0
        run.step(1)
        n = run.particles.np
        for f in FAMILIES:
            k, t = ctx.kernel_stats(f)
            e = fam.setdefault(f, {"launches": 0, "ms": 0.0, "pairs": 0.0, "n": []})
            e["launches"] += k - marks[f][0]
            e["ms"] += t - marks[f][1]
            if f in SELF_FAMILIES and k > marks[f][0]:
                e["pairs"] += (k - marks[f][0]) * float(n) * float(n)
                e["n"].append(n)
        if clo:
            e = fam["closure_near"]
            e["pairs"] += run.particles.closure_stats()["near_pairs"] - near0
            e["n"].append(n)
    ctx.set_timing(False)
    return fam


def run_config(hybrid, particles, steps, warmup, precision, stretch=True, closure=None,
               mlp=True, perturb=0.0, families=True):
    run, case = prepare(hybrid, particles, warmup + steps + 3, perturb)
    if precision != "fp64":
        run.set_precision(precision)
    if closure is not None:
        enable_closure(run.particles, closure, mlp)
    run.set_flags(True, stretch)
    run.step(warmup)
    out = timed(run, steps)
    if families:
        out["families"] = kernel_families(run, 1 if particles > 50000 else 2)
    if closure is not None:
        out["closure"] = run.particles.closure_stats()
    run.close()
    return out


def enable_closure(field, closure, mlp=True):
    from fluxvortex_amd.mlp import ClosureMLP, closure_params
    net = ClosureMLP(closure_params(seed=0)) if mlp else None
    clo = dict(closure)
    kl = clo.pop("kappa_local", 0.0)
    field.enable_closure(net, **clo)
    if kl:
        field.set_closure_local(kl)
    if mlp:
        field.closure_train(rounds=800)


def cl_err(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def accuracy(hybrid, particles, steps, warmup, closure=None):
    """CL over the timed window (warmup + steps) from bitwise-identical states:
    FP64 against FP64 with alpha x (1 + 1e-14) (the step's own sensitivity),
    mixed (and mixed + learned closure) against FP64; for the stretched window
    (the timed workload) and the unstretched free wake (the reference
    HybridSolver's own advect)."""
    out = {}
    for stretch in (True, False):
        kw = dict(stretch=stretch, families=False)
        n = warmup + steps
        ref = run_config(hybrid, particles, n, 0, "fp64", **kw)["cl"]
        per = run_config(hybrid, particles, n, 0, "fp64", perturb=1e-14, **kw)["cl"]
        mix = run_config(hybrid, particles, n, 0, "mixed", **kw)["cl"]
        sens = np.abs(per - ref) / np.max(np.abs(ref))
        e = {"steps": n, "fp64_self_sensitivity": float(sens.max()),
             "mixed_cl_max_rel_err_vs_fp64": cl_err(mix, ref)}
        # steps in which FP64's own response stays below 1e-6: there the
        # comparison measures arithmetic, beyond it the dynamics
        well = int(np.argmax(sens > 1e-6)) if np.any(sens > 1e-6) else n
        e["well_conditioned_steps"] = well
        if well:
            e["mixed_cl_max_rel_err_well_conditioned"] = cl_err(mix[:well], ref[:well])
        if closure is not None:
            ml = run_config(hybrid, particles, n, 0, "mixed", closure=closure, **kw)
            e["mixed_learned_cl_max_rel_err_vs_fp64"] = cl_err(ml["cl"], ref)
            if well:
                e["mixed_learned_cl_max_rel_err_well_conditioned"] = cl_err(ml["cl"][:well],
                                                                            ref[:well])
            e["closure"] = ml["closure"]
        out["stretched" if stretch else "unstretched"] = e
    return out


def tuned_shape(ctx, family, n):
    """(lanes, targets per lane, workgroups per CU) the autotuner chose for a
    self sum of n particles (the library's default shape if untuned)."""
    lg = int(np.ceil(np.log2(max(n, 1))))   # the tuner's size class (fv_biot_savart.hip)
    for e in ctx.autotune_report():
        if e[0] == family and e[1] == lg and e[2] == lg:
            return e[3], e[4], e[5]
    return (256, 4, 16) if n >= 65536 else (64, 2 if n >= 16384 else 1, 16)


def pair_split(n, lanes, tpt, per_cu, cus=256):
    """Source split of one pair-sum launch (fv_biot_savart.hip plan_pairs_shape)."""
    gx = -(-n // (lanes * tpt))
    tiles = -(-n // lanes)
    split = max(1, min(tiles, -(-(per_cu * cus) // gx), 65535))
    chunk = -(-tiles // split) * lanes
    return -(-n // chunk)


def roofline(fam, ctx):
    """Dominant particle kernel: algorithmic flops per launch / average launch
    time (HIP events), FP64 VALU peak.  traffic: HBM bytes per launch from the
    counter calibration (fetch per particle) and the exact partial-sum writes."""
    cands = {k: v for k, v in fam.items() if k in SELF_FAMILIES and v["launches"]}
    name = max(cands, key=lambda k: cands[k]["ms"])
    e = cands[name]
    avg_s = e["ms"] / e["launches"] / 1e3
    flops = FLOPS[name] * e["pairs"] / e["launches"]
    achieved = flops / avg_s
    n = float(np.mean(e["n"]))
    lanes, tpt, per_cu = tuned_shape(ctx, name, int(n))
    split = pair_split(int(n), lanes, tpt, per_cu)
    traffic = FETCH_BYTES_PER_PARTICLE * n + split * n * 3 * 8
    return name, {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12,
                  "unit": "TFLOP/s", "frac": achieved / PEAK_FP64, "traffic": traffic,
                  "kernel": f"k_pairs<PartOp> ({name})", "avg_launch_ms": avg_s * 1e3,
                  "particles_per_launch": n, "pairs_per_launch": e["pairs"] / e["launches"],
                  "flops_per_pair": FLOPS[name], "flops_per_launch": flops,
                  "shape": {"lanes": lanes, "targets_per_lane": tpt, "workgroups_per_cu": per_cu,
                            "source_split": split},
                  "algorithmic_bytes": n * (56 + 48 + 24),
                  "note": "FP64 VALU pair sum; MI355X FP64 vector peak 78.6 TFLOP/s; traffic "
                          "= calibrated fetch per particle x N + split x N x 24 B partial sums"}


def roofline_others(fam, P=120, rings_per_convert=12, mixed=False):
    """The other kernel families of the hybrid step against their own roofs:
    AIC assembly and conversion as HBM GB/s, the dense solve against the FP64
    (FP32 for mixed) MFMA peak."""
    out = {}

    def per(name):
        e = fam.get(name)
        return (e["ms"] / e["launches"] / 1e3) if e and e["launches"] else None

    t = per("aic")
    if t:
        b = 8.0 * P * P + 8.0 * 15 * P
        out["aic"] = {"bound": "hbm", "bytes_per_launch": b, "avg_launch_us": t * 1e6,
                      "achieved_GBps": b / t / 1e9, "peak_GBps": PEAK_HBM / 1e9,
                      "frac": b / t / PEAK_HBM}
    t = per("convert")
    if t:
        b = rings_per_convert * (104.0 + 4 * 56.0)
        out["convert"] = {"bound": "hbm", "bytes_per_launch": b, "avg_launch_us": t * 1e6,
                          "achieved_GBps": b / t / 1e9, "peak_GBps": PEAK_HBM / 1e9,
                          "frac": b / t / PEAK_HBM}
    t = per("dense_solve")
    if t:
        f = 2.0 / 3.0 * P ** 3 + 2.0 * P * P
        peak = PEAK_FP32_MFMA if mixed else PEAK_FP64
        out["dense_solve"] = {"bound": "mfma", "n": P, "flops_per_launch": f,
                              "avg_launch_us": t * 1e6, "achieved_TFLOPs": f / t / 1e12,
                              "peak_TFLOPs": peak / 1e12, "frac": f / t / peak}
    return out


def mlp_roofline(nb=1 << 18):
    """The closure network's forward pass on nb samples (BF16 MFMA, FP32
    accumulation) against the dense BF16 peak, kernel time from HIP events."""
    from fluxvortex_amd import _lib
    from fluxvortex_amd.mlp import ClosureMLP
    ctx = _lib.default_context()
    net = ClosureMLP(seed=3)
    X = np.random.default_rng(0).uniform(-1, 1, (nb, 31)).astype(np.float32)
    net.forward(X[:1024])
    ctx.set_timing(True)
    for _ in range(3):
        net.forward(X)
    k, ms = ctx.kernel_stats("mlp_forward")
    T = np.random.default_rng(1).uniform(-1, 1, (nb, 3)).astype(np.float32)
    for _ in range(3):
        net.gradients(X, T, np.ones(3, np.float32))
    kt, mst = ctx.kernel_stats("mlp_train")
    ctx.set_timing(False)
    t = ms / k / 1e3
    f = MLP_FLOPS_PER_SAMPLE * nb
    out = {"bound": "mfma", "samples": nb, "flops_per_sample": MLP_FLOPS_PER_SAMPLE,
           "forward_avg_launch_us": t * 1e6, "forward_TFLOPs": f / t / 1e12,
           "peak_TFLOPs": PEAK_BF16 / 1e12, "frac": f / t / PEAK_BF16}
    if kt:
        tt = mst / kt / 1e3
        out["gradients_avg_launch_us"] = tt * 1e6
        out["gradients_TFLOPs"] = 3 * f / tt / 1e12   # forward + backward + weight gradient
    return out


def host_threads():
    """Host threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when that is set (the GPU box's share)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n or 1)


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(state, threads, reps=5):
    """The oracle's particle sums of one hybrid step on the very state the GPU
    line timed (3 Winckelmans velocity + 3 Winckelmans Jacobian + 1 Gaussian
    Jacobian under the clamp rule), all targets x all sources, each kind the
    median of `reps` runs on all host threads."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    pos, al, sg = state["pos"], state["gamma"], state["sigma"]
    n = pos.shape[0]
    oracle.particle_vel(pos[:16], pos, al, sg, "winckelmans", threads)   # warm-up

    def med(fn, *args, **kw):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn(*args, **kw)
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), sum(ts)

    t_vel, s1 = med(oracle.particle_vel, pos, pos, al, sg, "winckelmans", threads)
    t_jw, s2 = med(oracle.particle_jac, pos, pos, al, sg, "winckelmans", threads)
    t_jg, s3 = med(oracle.particle_jac, pos, pos, al, sg, "gaussianerf", threads, rule=1)
    t_step = 3 * t_vel + 3 * t_jw + t_jg
    return {"value": 1.0 / t_step, "unit": "steps/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "particles": n,
            "sample": (f"oracle particle sums of one step on the timed state ({n} particles): "
                       f"3 Winckelmans velocity + 3 Winckelmans Jacobian + 1 Gaussian Jacobian, "
                       f"all targets x all sources, each kind the median of {reps} runs; "
                       f"ring/UVLM part (<0.1%) omitted"),
            "sample_seconds": s1 + s2 + s3}


def published_sizes():
    """The reference's only published timings: the collapsed ring routine at
    (N points, M rings) = (500, 2000), (1000, 5000) and 10 000 x 10 000
    (README.md:178-180), plus its benchmark's third size (500, 10 000)
    (tests/test_benchmark.py:20), through the patch surface with the host
    copies included, seeded as that benchmark; median of 7 after a warm-up."""
    from fluxvortex_amd import ptera_patch
    ns = ptera_patch.default_target()
    ptera_patch.patch(verbose=False)
    try:
        fn = ns.collapsed_velocities_from_ring_vortices
        rng = np.random.default_rng(42)
        pub = {(500, 2000): (43.0, 18.0), (1000, 5000): (190.0, 83.0),
               (10000, 10000): (3800.0, 350.0), (500, 10000): (None, None)}
        out = []
        for (N, M), (cpu_ms, gpu_ms) in pub.items():
            points = rng.standard_normal((N, 3))
            br, fr, fl, bl = (rng.standard_normal((M, 3)) for _ in range(4))
            g = rng.standard_normal(M) * 0.5
            rc = np.full(M, 0.03)
            sing = np.zeros(4, dtype=np.int64)
            ages = rng.uniform(0, 1, M)
            args = (points, br, fr, fl, bl, g, rc, sing, ages, 1.5e-5)
            fn(*args)
            ts = []
            for _ in range(7):
                t0 = time.perf_counter()
                fn(*args)
                ts.append(time.perf_counter() - t0)
            ms = float(np.median(ts)) * 1e3
            e = {"N": N, "M": M, "ms_host_copies_included": ms,
                 "gpairs_per_s": 4.0 * N * M / (ms / 1e3) / 1e9}
            if gpu_ms:
                e.update({"reference_gpu_ms": gpu_ms, "reference_cpu_ms": cpu_ms,
                          "speedup_vs_reference_gpu": gpu_ms / ms})
            out.append(e)
        return out
    finally:
        ptera_patch.unpatch(verbose=False)


def static_wing_block():
    """C1: the static NACA 2412 wing (7 x 18 cosine panels, alpha 5 deg,
    prescribed ring wake over 6 chords), half wing (126 panels) and mirrored
    (252 panels): whole run steps/s, final CL / CDi beside PteraSoftware's
    published 0.5180 / 0.0175 (README.md:44)."""
    from fluxvortex_amd import hybrid
    out = {}
    for sym, P in ((False, 126), (True, 252)):
        case = hybrid.static_wing_case(symmetric=sym)
        run = hybrid.HybridRun(case)
        run.step(1)
        run._ctx.synchronize()
        ctx = run._ctx
        ctx.event_record(2)
        run.step(case.num_steps - 1)
        ctx.event_record(3)
        ms = ctx.event_elapsed_ms(2, 3)
        cl, cdi, _ = run.read()
        run.close()
        out[f"P{P}"] = {"panels": P, "steps": case.num_steps - 1,
                        "steps_per_s": (case.num_steps - 1) / (ms / 1e3),
                        "cl": float(cl[-1]), "cdi": float(cdi[-1])}
    out["published_ptera_cl_cdi"] = [0.5180, 0.0175]
    return out


def raw_kernel_sweep(oracle_threads):
    """C3, the seeded synthetic raw-kernel sweep at all six sizes: M rings on
    the perturbed sheet (Gamma ~ N(0,1), rc0 = 0.03, age ~ U(0,1)), N
    Winckelmans particles in the 20 x 10 x 2 slab (alpha ~ N(0, 0.01^2),
    sigma = 0.05); targets = the ring centroids + every particle.  Velocity
    from rings + particles at every target and the particle Jacobian at every
    particle (stretching S = J alpha), FP64 and mixed: kernel time from HIP
    events -> G pairs/s; mixed vs FP64 per-target relative errors (histogram
    over every target); FP64 vs the oracle on 256 sampled targets.  The
    Gaussian kernel is run at the largest size."""
    sys.path.insert(0, str(ROOT / "tests"))
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle
    from helpers import sheet_rings
    from fluxvortex_amd import _lib, kernel as K, vortex_kernels as vk
    ctx = _lib.default_context()
    bins = [0.0] + [10.0 ** e for e in range(-16, 0)] + [np.inf]

    def hist(err):
        h, _ = np.histogram(err, bins=bins)
        return {"edges": ["0"] + [f"1e{e}" for e in range(-16, 0)] + ["inf"],
                "counts": h.tolist(), "p50": float(np.median(err)),
                "p99": float(np.quantile(err, 0.99)), "max": float(np.max(err))}

    def per_target(a, b):
        nb = np.linalg.norm(b.reshape(len(b), -1), axis=1)
        d = np.linalg.norm((a - b).reshape(len(b), -1), axis=1)
        return d / np.maximum(nb, 1e-300)

    def kern(fn, family):
        ctx.set_timing(True)
        t0 = time.perf_counter()
        r = fn()
        wall = time.perf_counter() - t0
        k, ms = ctx.kernel_stats(family)
        ctx.set_timing(False)
        return r, ms, wall * 1e3

    out = []
    sizes = [(m, n) for m in (1000, 10000) for n in (10000, 100000, 1000000)]
    for m_rings, n_part in sizes:
        for kernel in (("winckelmans", "gaussianerf") if (m_rings, n_part) == sizes[-1]
                       else ("winckelmans",)):
            nx = 40 if m_rings == 1000 else 100
            br, fr, fl, bl, g, rc, age = sheet_rings(nx, m_rings // nx, seed=42)
            rng = np.random.default_rng(42)
            pos = rng.uniform([0.0, -5.0, -1.0], [20.0, 5.0, 1.0], (n_part, 3))
            al = rng.normal(0.0, 0.01, (n_part, 3))
            sg = np.full(n_part, 0.05)
            tgt = np.concatenate([0.25 * (br + fr + fl + bl), pos])
            s4 = np.zeros(4, np.int64)
            res = {"M": m_rings, "N": n_part, "kernel": kernel, "targets": len(tgt)}
            vals = {}
            for prec in ("fp64", "mixed"):
                vk.set_precision(prec)
                try:
                    ur, ms_r, _ = kern(lambda: vk.collapsed_velocities_from_ring_vortices(
                        tgt, br, fr, fl, bl, g, rc, s4, age, 1.5e-5), "seg_pairs")
                finally:
                    vk.set_precision("fp64")
                up, ms_p, _ = kern(lambda: K.velocity_from_particles(
                    tgt, pos, al, sg, kernel=kernel, precision=prec), "particle_vel")
                J, ms_j, _ = kern(lambda: K.jacobian_from_particles(
                    pos, al, pos, al, sg, kernel=kernel, precision=prec), "particle_jac")
                S = np.einsum("tij,tj->ti", J, al)
                vals[prec] = (ur + up, S)
                seg_pairs = 4.0 * m_rings * len(tgt)
                par_pairs = float(n_part) * len(tgt)
                jac_pairs = float(n_part) * n_part
                res[prec] = {"velocity_kernel_ms": ms_r + ms_p, "jacobian_kernel_ms": ms_j,
                             "segment_gpairs_per_s": seg_pairs / (ms_r / 1e3) / 1e9,
                             "particle_velocity_gpairs_per_s": par_pairs / (ms_p / 1e3) / 1e9,
                             "particle_jacobian_gpairs_per_s": jac_pairs / (ms_j / 1e3) / 1e9,
                             "gpairs_per_s": (seg_pairs + par_pairs + jac_pairs)
                             / ((ms_r + ms_p + ms_j) / 1e3) / 1e9}
            u64, S64 = vals["fp64"]
            umx, Smx = vals["mixed"]
            res["mixed_vs_fp64"] = {
                "velocity_normwise": float(np.max(np.abs(umx - u64)) / np.max(np.abs(u64))),
                "stretching_normwise": float(np.max(np.abs(Smx - S64)) / np.max(np.abs(S64))),
                "velocity_per_target": hist(per_target(umx, u64)),
                "stretching_per_target": hist(per_target(Smx, S64))}
            pick = np.sort(rng.choice(len(tgt), 256, replace=False))
            ppick = np.sort(rng.choice(n_part, 256, replace=False))
            ref_u = (oracle.collapsed(tgt[pick], br, fr, fl, bl, g, rc, age, 1.5e-5,
                                      threads=oracle_threads)
                     + oracle.particle_vel(tgt[pick], pos, al, sg, kernel, oracle_threads))
            ref_S = np.einsum("tij,tj->ti", oracle.particle_jac(pos[ppick], pos, al, sg, kernel,
                                                                oracle_threads), al[ppick])
            res["fp64_vs_oracle_256"] = {
                "velocity_normwise": float(np.max(np.abs(u64[pick] - ref_u))
                                           / np.max(np.abs(ref_u))),
                "stretching_normwise": float(np.max(np.abs(S64[ppick] - ref_S))
                                             / np.max(np.abs(ref_S))),
                "velocity_per_target": hist(per_target(u64[pick], ref_u)),
                "stretching_per_target": hist(per_target(S64[ppick], ref_S))}
            log(f"[bench] C3 M={m_rings} N={n_part} {kernel}: FP64 "
                f"{res['fp64']['gpairs_per_s']:.0f} G pairs/s")
            out.append(res)
    return out


def flapping_closure_block(closure_candidates):
    """C5: the 192-panel flapping airplane (three flapping periods of free,
    unstretched particle wake, then extended to 100 000 particles by shifted
    copies of the oldest period): FP64 direct sums against mixed + learned
    closure (the cheapest passing candidate), BIG_STEPS + 1 free-wake steps
    each (the closure's check falls in the window), CL error vs FP64."""
    from fluxvortex_amd import hybrid

    def at(n_particles, prefill=129):
        case = hybrid.flapping_case(num_steps=prefill + BIG_STEPS + 8, free_wake=True,
                                    stretch=False, max_particles=n_particles + 4096)
        run = hybrid.HybridRun(case)
        run.set_flags(True, False)
        run.step(prefill)
        st = run.particles.state_dict()
        n0 = st["pos"].shape[0]
        block = min(n0, 43 * 128)
        parts = {k: [v] for k, v in st.items()}
        need, k = n_particles - n0, 1
        while need > 0:
            take = min(need, block)
            p = st["pos"][:take].copy()
            p[:, 0] += k * 10.0 * 1.0          # V T
            parts["pos"].append(p)
            parts["gamma"].append(st["gamma"][:take])
            parts["sigma"].append(st["sigma"][:take])
            parts["age"].append(st["age"][:take] + k * 1.0)
            need -= take
            k += 1
        run.particles.load_state_dict({k: np.concatenate(v) for k, v in parts.items()})
        return run

    def window(run):
        run.step(1)
        t = timed(run, BIG_STEPS)
        run.close()
        return t

    n = 100_000
    r64 = window(at(n))
    out = {"particles": r64["np"], "timed_steps": BIG_STEPS,
           "fp64_steps_per_s": BIG_STEPS / (r64["ms"] / 1e3)}
    for clo in closure_candidates:
        run = at(n)
        run.set_precision("mixed")
        enable_closure(run.particles, clo)
        t = window_keep(run)
        st = t.pop("closure")
        if st["active"] and st["rejections"] == 0:
            out.update({"closure_params": clo, "closure": st,
                        "mixed_learned_steps_per_s": BIG_STEPS / (t["ms"] / 1e3),
                        "mixed_learned_speedup_vs_fp64": r64["ms"] / t["ms"],
                        "mixed_learned_cl_max_rel_err_vs_fp64": cl_err(t["cl"], r64["cl"])})
            break
    return out


def window_keep(run):
    run.step(1)
    t = timed(run, BIG_STEPS)
    t["closure"] = run.particles.closure_stats()
    run.close()
    return t


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


def near_roofline(fam):
    """The closure's exact near field (k_near) over a timed step: the pairs it
    evaluated (device count) per second, and as FP64 VALU flops against the
    78.6 TFLOP/s peak with the direct sums' algorithmic flops per pair
    (velocity 31, stretching 66, Gaussian vorticity 72; 3 + 3 + 1 passes per
    step, mixed precision: FP32 kernel factors, FP64 geometry and sums)."""
    e = fam.get("closure_near")
    if not e or not e["launches"] or not e["pairs"]:
        return None
    sec = e["ms"] / 1e3
    flops_per_pair = (3 * 31 + 3 * 66 + 72) / 7.0
    ach = e["pairs"] * flops_per_pair / sec
    return {"bound": "valu", "pairs": e["pairs"], "launches": e["launches"],
            "avg_launch_us": e["ms"] / e["launches"] * 1e3, "gpairs_per_s": e["pairs"] / sec / 1e9,
            "achieved_TFLOPs": ach / 1e12, "peak_TFLOPs": PEAK_FP64 / 1e12,
            "frac": ach / PEAK_FP64}


def plunging_sweep_block():
    """The reference's reduced-frequency sweep (k = 0.5 / 0.2 / 0.1 x N_keep =
    10 / 20 / 0, free particle wake, 150 steps each): the nine runs one after
    another (eager, the default shapes), then as one batch in graph mode (one
    graph launch per time step once every run is steady).  Reports both wall
    times and every run's CL amplitude over Theodorsen's."""
    from fluxvortex_amd import hybrid
    cases = hybrid.plunging_sweep_cases(ks=(0.5, 0.2, 0.1), n_keeps=(10, 20, 0))
    steps = cases[0].num_steps
    hybrid.run_case(cases[0], steps=12)   # loads the kernels
    t0 = time.perf_counter()
    seq = [hybrid.run_case(c) for c in cases]
    t_seq = time.perf_counter() - t0
    b = hybrid.HybridBatch(cases)
    try:
        t0 = time.perf_counter()
        b.step(steps)
        outs = [r.outputs() for r in b.runs]
        t_bat = time.perf_counter() - t0
        launches, eager = b.stats()
    finally:
        b.close()
    runs = []
    for c, o, q in zip(cases, outs, seq):
        runs.append({"k": c.meta["k"], "n_keep": c.n_keep, "particles": o["np"],
                     "theodorsen_ratio": hybrid.theodorsen_ratio(c, o["cl"]),
                     "theodorsen_ratio_eager_default": hybrid.theodorsen_ratio(c, q["cl"])})
    return {"runs": len(cases), "steps": steps, "sequential_eager_s": t_seq,
            "batched_graph_s": t_bat, "speedup": t_seq / t_bat,
            "graph_launches": launches, "eager_batch_steps": eager,
            "run_steps_per_s_batched": len(cases) * steps / t_bat, "cases": runs}


# (cell edge h in m, cutoff kappa); kappa >= 2 keeps every neighbouring cell
# (centre distance up to sqrt(3) h) in the exact near field
CLOSURE_CANDIDATES = ((1.25, 2.0), (1.5, 2.0), (2.0, 2.0), (2.5, 2.0), (2.0, 3.0))


def closure_params(cell, kappa):
    return dict(cell=cell, kappa=kappa, check_every=10, tol=1e-3, train_samples=4096,
                train_iters=2, lr=2e-3, kappa_local=2.0 * kappa)


def big_block(hybrid, n, skip_mixed, candidates=CLOSURE_CANDIDATES):
    """The 100k block: FP64, mixed, then the closure candidates cheapest first
    with the network and without it (expansion only), each kept only if every
    one of its own checks passed; accuracy windows as for the 20k line."""
    big = {"timed_steps": BIG_STEPS}
    log(f"[bench] {n} particles, FP64")
    b64 = run_config(hybrid, n, BIG_STEPS, 1, "fp64")
    big["particles"] = b64["np"]
    big["fp64_steps_per_s"] = BIG_STEPS / (b64["ms"] / 1e3)
    big["fp64_gpairs_per_s"] = (b64["seg_pairs"] + b64["particle_pairs"]) / (b64["ms"] / 1e3) / 1e9
    if skip_mixed:
        return big, b64, None
    log(f"[bench] {n} particles, mixed")
    bmx = run_config(hybrid, n, BIG_STEPS, 1, "mixed")
    big["mixed_steps_per_s"] = BIG_STEPS / (bmx["ms"] / 1e3)
    # every candidate with and without the network; of those whose closure
    # passed every one of its own checks (a rejected check sends the steps back
    # to the direct sums), the fastest is kept
    chosen = {}
    for mlp in (True, False):
        tried = []
        for cell, kappa in candidates:
            clo = closure_params(cell, kappa)
            log(f"[bench] {n} particles, mixed + closure (network {mlp}), cell {cell} "
                f"kappa {kappa}")
            r = run_config(hybrid, n, BIG_STEPS, 1, "mixed", closure=clo, mlp=mlp,
                           families=False)
            st = r["closure"]
            tried.append({"cell": cell, "kappa": kappa, "checks": st["checks"],
                          "rejections": st["rejections"], "last_err": st["last_err"],
                          "steps_per_s": BIG_STEPS / (r["ms"] / 1e3)})
            ok = st["active"] and st["rejections"] == 0
            if ok and (mlp not in chosen or r["ms"] < chosen[mlp][1]["ms"]):
                chosen[mlp] = (clo, r)
        big["closure_candidates" if mlp else "expansion_only_candidates"] = tried
    if True in chosen:   # kernel families of the kept learned configuration
        clo = chosen[True][0]
        chosen[True] = (clo, run_config(hybrid, n, BIG_STEPS, 1, "mixed", closure=clo))
    if True in chosen:
        clo, bml = chosen[True]
        big["mixed_learned_steps_per_s"] = BIG_STEPS / (bml["ms"] / 1e3)
        big["mixed_learned_speedup_vs_fp64"] = b64["ms"] / bml["ms"]
        big["closure"] = bml["closure"]
        big["closure_params"] = clo
    if False in chosen:
        clo, bex = chosen[False]
        big["expansion_only_steps_per_s"] = BIG_STEPS / (bex["ms"] / 1e3)
        big["expansion_only_params"] = clo
        if True in chosen:
            big["network_speedup_at_equal_tolerance"] = bex["ms"] / chosen[True][1]["ms"]
    return big, b64, chosen.get(True)


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
    ap.add_argument("--closure", default=None,
                    help="CELL,KAPPA: the 100k block's only closure candidate (default: all)")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip C1 / C3 / C5, the published sizes and the MLP roofline")
    a = ap.parse_args(argv)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": a.gpus}))
        return 2
    from fluxvortex_amd import _lib, hybrid
    _lib.load_library()
    ctx = _lib.default_context()
    # tile shapes of the pair sums (and the closure's near field) timed at
    # first use, for every configuration below alike
    ctx.set_autotune(not a.no_autotune)
    log(f"[bench] heaving wing, {a.particles} particles, FP64, warmup {a.warmup}, steps {a.steps}")
    run, case = prepare(hybrid, a.particles, a.warmup + a.steps + 3)
    run.set_flags(True, True)
    run.step(a.warmup)
    state_timed = run.particles.state_dict()
    m64 = timed(run, a.steps)
    m64["families"] = kernel_families(run, 2)
    run.close()
    sec = m64["ms"] / 1e3
    pairs_total = m64["seg_pairs"] + m64["particle_pairs"]
    name, roof = roofline(m64["families"], ctx)
    out = {
        "metric": "hybrid UVLM-VPM steps/s & G pair-interactions/s: FP64 vs mixed+learned",
        "value": a.steps / sec, "unit": "steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": m64["ms"] / a.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "heaving-wing hybrid UVLM-VPM step: 120 panels, N_keep=10, "
                               "Winckelmans sigma=0.5 V dt, free wake + rVPM stretching + "
                               "relaxation (reference CPU pair rule), RK3",
                   "particles": m64["np"], "panels": 120, "precision": "fp64",
                   "parallelism": "single GPU"},
        "gpairs_per_s": pairs_total / sec / 1e9,
        "segment_gpairs_per_s": m64["seg_pairs"] / sec / 1e9,
        "particle_gpairs_per_s": m64["particle_pairs"] / sec / 1e9,
        "roofline": roof,
        "roofline_others": roofline_others(m64["families"]),
        "kernel_family_ms_per_launch": {k: round(v["ms"] / v["launches"], 4)
                                        for k, v in m64["families"].items() if v["launches"]},
    }
    if not a.no_mixed:
        log("[bench] same workload, mixed precision")
        mmx = run_config(hybrid, a.particles, a.steps, a.warmup, "mixed")
        out["mixed"] = {"steps_per_s": a.steps / (mmx["ms"] / 1e3),
                        "speedup_vs_fp64": m64["ms"] / mmx["ms"],
                        "roofline_others": roofline_others(mmx["families"], mixed=True)}
        log("[bench] accuracy windows (FP64 self-sensitivity, mixed)")
        acc = accuracy(hybrid, a.particles, a.steps, a.warmup)
        out["mixed"]["accuracy"] = acc
        out["mixed"]["cl_max_rel_err_vs_fp64"] = acc["unstretched"]["mixed_cl_max_rel_err_vs_fp64"]
        out["mixed"]["cl_error_window"] = "unstretched free wake (HybridSolver's advect)"
    if a.big:
        cands = CLOSURE_CANDIDATES
        if a.closure:
            cands = (tuple(float(v) for v in a.closure.split(",")),)
        big, b64, chosen = big_block(hybrid, a.big, a.no_mixed, cands)
        if chosen is not None:
            clo, bml = chosen
            if "families" in bml:
                big["mlp_roofline_in_step"] = {
                    k: bml["families"][k] for k in ("mlp_forward", "mlp_train", "closure_near",
                                                    "closure_far", "closure_residual")
                    if k in bml["families"]}
                big["closure_near_roofline"] = near_roofline(bml["families"])
            log("[bench] 100k accuracy windows (FP64 self-sensitivity, mixed, mixed + learned)")
            acc = accuracy(hybrid, a.big, BIG_STEPS, 1, closure=clo)
            big["accuracy"] = acc
            big["mixed_learned_cl_max_rel_err_vs_fp64"] = \
                acc["unstretched"]["mixed_learned_cl_max_rel_err_vs_fp64"]
        out["heaving_100k"] = big
    if not a.no_extras:
        log("[bench] MLP roofline")
        out["mlp_roofline"] = mlp_roofline()
        log("[bench] published collapsed-ring sizes through the patch surface")
        out["published_sizes"] = published_sizes()
        log("[bench] C1 static wing")
        out["static_wing"] = static_wing_block()
        log("[bench] batched reduced-frequency sweep (9 runs, graph mode)")
        out["plunging_sweep"] = plunging_sweep_block()
        if not a.no_mixed:
            log("[bench] C5 flapping airplane at 100k with the closure")
            out["flapping_100k"] = flapping_closure_block(
                [closure_params(c, k) for c, k in CLOSURE_CANDIDATES])
        log("[bench] C3 raw-kernel sweep")
        out["raw_kernel_sweep"] = raw_kernel_sweep(host_threads())
    if not a.no_flutter:
        log("[bench] batched Goland flutter sweep, 64 speeds")
        out["flutter_sweep"] = flutter_block()
    out["autotune"] = [dict(zip(("family", "log2_targets", "log2_sources", "lanes",
                                 "targets_per_lane", "workgroups_per_cu", "best_ms"), e))
                       for e in ctx.autotune_report() if e[1] >= 14]
    if not a.no_cpu_baseline:
        threads = host_threads()
        log(f"[bench] CPU baseline (oracle, {threads} threads, {a.particles} particles)")
        out["cpu_baseline"] = cpu_baseline(state_timed, threads)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
