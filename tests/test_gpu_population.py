"""Device population control and circulation ledger (fv_vpm_drop_oldest,
fv_vpm_merge_host, fv_vpm_ledger_host, fv_hybrid_set_population) against the
oracle: the device and the CPU statement pick the same pairs, so the merged
fields agree bit for bit."""
import numpy as np
import pytest

from fluxvortex_amd import hybrid
from fluxvortex_amd.particles import VortexParticleField

pytestmark = pytest.mark.gpu

REF = np.array([[0.5 + 0.2 * k, -1.0 + 0.25 * k, 0.0] for k in range(9)])


def _cloud(n, seed, aligned=0.02, spread=(30, 3, 0.5)):
    rng = np.random.default_rng(seed)
    pos = rng.uniform([0, -spread[1], -spread[2]], spread, (n, 3))
    al = rng.normal(0, 0.01, (n, 3))
    al[:, 1] += aligned
    sg = rng.uniform(0.05, 0.15, n)
    return pos, al, sg


def _field(pos, al, sg, cap=None):
    f = VortexParticleField(max_particles=cap or len(sg) + 16, kernel="winckelmans")
    f.add_particles_batch(pos, al, sg)
    return f


def _state(f):
    return f.positions.copy(), f.gammas.copy(), f.sigmas.copy(), f._age.copy()


def test_drop_matches_oracle(oracle):
    pos, al, sg = _cloud(5000, 1)
    f = _field(pos, al, sg)
    ages = np.zeros(len(sg))
    f.drop_oldest(3000)
    p, a, s, g = _state(f)
    rp, ra, rs, rg = oracle.pop_drop(pos, al, sg, ages, 3000)
    assert f.np == 3000
    assert np.array_equal(p, rp) and np.array_equal(a, ra) and np.array_equal(s, rs)
    f.drop_oldest(4000)   # below the count: nothing
    assert f.np == 3000


@pytest.mark.parametrize("case", ["mixed", "one_cell", "misaligned", "protect_dist", "strict"])
def test_merge_matches_oracle(oracle, case):
    n = 6000
    kw = dict(eps=1e-3, vinf_norm=10.0, protect=64, protect_dist=0.0)
    if case == "one_cell":
        pos, al, sg = _cloud(n, 2, spread=(20.5, 0.3, 0.05))
        pos[:, 0] += 20.0
    elif case == "misaligned":
        pos, al, sg = _cloud(n, 3, aligned=0.0)
    elif case == "protect_dist":
        pos, al, sg = _cloud(n, 4)
        kw["protect_dist"] = 8.0
    elif case == "strict":
        pos, al, sg = _cloud(n, 5)
        kw["eps"] = 1e-9
    else:
        pos, al, sg = _cloud(n, 6)
    print("merge case", case, flush=True)
    f = _field(pos, al, sg)
    ages = np.zeros(n)
    nm = f.merge(REF, **kw)
    rp, ra, rs, rg, rnm = oracle.pop_merge(pos, al, sg, ages, REF, **kw)
    assert nm == rnm and f.np == len(rs)
    p, a, s, _ = _state(f)
    assert np.array_equal(p, rp) and np.array_equal(a, ra) and np.array_equal(s, rs)
    if case == "strict":   # only near-coincident pairs pass a 1e-9 bar
        assert nm < 50
    elif case != "misaligned":
        assert nm > 50


def test_ledger_matches_oracle(oracle):
    pos, al, sg = _cloud(20000, 7)
    f = _field(pos, al, sg)
    got = f.circulation_ledger()
    ref = oracle.ledger(al)
    assert np.abs(got - ref).max() < 1e-13 * np.abs(al).sum()
    before = got.copy()
    f.merge(REF, eps=1e-2, vinf_norm=10.0)
    assert np.abs(f.circulation_ledger() - before).max() < 1e-12 * np.abs(al).sum()


def test_hybrid_population_control():
    """Heaving wing with a prescribed particle wake and no relaxation (alpha is
    then constant).  'drop' with the cap at the unlimited run's count after 30
    steps: the first 31 CL values are those of the unlimited run bit for bit
    and the field ends at the cap.  'merge' shrinks the field and keeps the
    ledger equal to the unmerged run's."""
    def run(scheme, steps, **kw):
        case = hybrid.heaving_wing_case(num_steps=40, free_wake=False, stretch=False, rlxf=0.0)
        r = hybrid.HybridRun(case)
        r.set_population(scheme, **kw)
        r.step(steps)
        return r

    base = run("none", 40)
    cl_base = base.read()[0]
    cap = run("none", 31).particles.np
    drop = run("drop", 40, max_particles=cap)
    assert drop.particles.np == cap < base.particles.np
    assert np.array_equal(drop.read()[0][:31], cl_base[:31])
    mer = run("merge", 40, eps=1e-2, protect=64)
    assert mer.particles.np < base.particles.np
    lb, lm = base.circulation_ledger(), mer.circulation_ledger()
    scale = np.abs(base.particles.gammas).sum()
    assert np.abs(lm - lb).max() < 1e-12 * scale


def test_borrowed_particle_field_keeps_its_case_alive():
    """A hybrid case's particle field used after the last reference to the case
    is gone (the field keeps the case alive; it used to read freed memory)."""
    import gc
    p = hybrid.HybridRun(hybrid.heaving_wing_case(num_steps=5)).particles
    gc.collect()
    assert p.np == 0
