"""Graph mode of the hybrid step and the batched reduced-frequency sweep:
steady steps replay from hipGraphs (one graph launch per step for a whole
batch of runs), bit for bit equal to the same launches run eagerly, and
within the oracle's own sensitivity of the CPU oracle."""
import numpy as np
import pytest

from test_gpu_hybrid import compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hyb(ctx):
    from fluxvortex_amd import hybrid
    return hybrid


def _same(a, b, keys=("cl", "cdi", "gamma", "pos", "alpha", "sigma_p", "age")):
    assert a["np"] == b["np"]
    for k in keys:
        assert np.array_equal(a[k], b[k]), k


def _plunge(hyb, k=0.5, n_keep=10, steps=40, **kw):
    return hyb.plunging_sweep_cases(ks=(k,), n_keeps=(n_keep,), num_steps=steps,
                                    max_particles=3000, **kw)[0]


@pytest.mark.parametrize("stretch,kernel", [(False, "gaussianerf"), (True, "winckelmans")])
def test_graph_replay_equals_eager_launches(hyb, stretch, kernel):
    """Heaving wing, free particle wake: every steady step replays from a
    graph, and the result equals graph mode run eagerly bit for bit."""
    case = _plunge(hyb, stretch=stretch, kernel=kernel)
    print(f"[graph] {case.name} stretch={stretch}", flush=True)
    run = hyb.HybridRun(case, graph=True)
    try:
        run.step(case.num_steps)
        g = run.outputs(particles=True)
        launches, eager = run.graph_stats()
    finally:
        run.close()
    e = hyb.run_case(case, particles=True, graph=2)
    _same(g, e)
    # the wake fills at n_keep + 1 rows; one eager steady step sizes the
    # buffers; the last step does not shed
    assert launches == case.num_steps - (case.n_keep + 3)
    assert launches + eager == case.num_steps
    assert g["np"] > 1000


def test_graph_mode_matches_oracle(hyb, oracle):
    """Graph-mode launches (fixed shapes, device-sized source slices) against
    the oracle at 1e-10 on every step: free particle wake without relaxation,
    which keeps the step well conditioned (with relaxation the clamped
    coincident pairs let round-off of any summation order set their alpha)."""
    case = _plunge(hyb, steps=30, n_keep=6, rlxf=0.0)
    run = hyb.HybridRun(case, graph=True)
    try:
        run.step(case.num_steps)
        dev = run.outputs(particles=True)
        assert run.graph_stats()[0] > 0
    finally:
        run.close()
    compare(hyb, oracle, case, dev=dev)


def test_graph_recaptures_after_a_setting_changes(hyb):
    """Turning stretching on mid-run makes the graphs stale: the next steady
    step runs eagerly, then replays resume, and the run equals graph mode run
    eagerly with the same change."""
    case = _plunge(hyb, steps=40, kernel="winckelmans")
    outs = []
    for mode in (True, 2):
        run = hyb.HybridRun(case, graph=mode)
        try:
            run.step(20)
            run.set_flags(True, True)
            run.step(case.num_steps - 20)
            outs.append(run.outputs(particles=True))
            if mode is True:
                launches, eager = run.graph_stats()
        finally:
            run.close()
    _same(*outs)
    assert launches > 0 and eager > case.n_keep + 3


@pytest.mark.parametrize("rlxf", [0.3, 0.0])
def test_batched_reduced_frequency_sweep(hyb, oracle, rlxf):
    """k = 0.5 / 0.2 / 0.1 x N_keep = 10 / 0 (six runs, uneven wakes and
    particle counts) in one batch: once every run is steady, one graph launch
    per time step advances all six; each run equals the same case run alone
    in graph mode bit for bit, and (without relaxation, where the step is
    well conditioned) the oracle within 1e-10 at every step."""
    cases = hyb.plunging_sweep_cases(ks=(0.5, 0.2, 0.1), n_keeps=(10, 0), num_steps=30,
                                     max_particles=2000, rlxf=rlxf)
    print(f"[batch] {[c.name for c in cases]}", flush=True)
    b = hyb.HybridBatch(cases)
    try:
        b.step(30)
        launches, eager = b.stats()
        outs = [r.outputs(particles=True) for r in b.runs]
    finally:
        b.close()
    assert launches + eager == 30
    assert launches == 30 - (10 + 3)        # the N_keep = 10 runs settle last
    nps = [o["np"] for o in outs]
    assert len(set(nps)) > 1                 # uneven particle counts
    for case, out in zip(cases, outs):
        _same(out, hyb.run_case(case, particles=True, graph=True))
    if rlxf == 0.0:
        for case, out in zip(cases[::2], outs[::2]):
            compare(hyb, oracle, case, dev=out)
