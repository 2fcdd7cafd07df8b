"""GPU-resident Monte Carlo loop (device proposals, Metropolis and commit):
bookkeeping checks at any size, and the FP32-pair path's per-term error."""
import numpy as np
import pytest

from oracle import oracle as O
from plum_amd import engine as E
from plum_amd import systems

pytestmark = pytest.mark.gpu
TERMS = ("lj", "real", "kspace", "self", "dipole", "bond", "wall")


def _bond_lengths(xyz, ms):
    out = []
    for m in range(len(ms) - 1):
        a, b = ms[m], ms[m + 1]
        if b - a > 1:
            out.append(np.linalg.norm(np.diff(xyz[a:b], axis=0), axis=1))
    return np.concatenate(out)


@pytest.mark.parametrize("name,moves", [("small", 3000), ("c0", 3000), ("slab_small", 3000)])
def test_device_loop_bookkeeping(name, moves):
    sd = systems.config(name)
    eng = E.Engine(sd)
    st = eng.mc_run(moves, sd["move_prob"], move_size=sd["move_size"], seed=11,
                    phantom=sd.get("phantom", 0), grafted=sd.get("grafted", 0), bond_len=1.0)
    acc = sum(st.accepted[k] for k in range(5))
    att = sum(st.attempted[k] for k in range(5))
    print("%s: moves=%d attempted=%s accepted=%s (%.1f%%) %.0f moves/s" % (
        name, st.moves, list(st.attempted), list(st.accepted), 100.0 * acc / max(att, 1),
        st.moves / st.seconds), flush=True)
    assert st.moves == moves and acc > 0
    run = st.totals.as_dict()
    full = eng.energy_full()
    for k in TERMS:
        if k == "dipole":
            continue  # running dipole follows the reference's one-move lag
        assert abs(run[k] - full[k]) <= 1e-9 * max(1.0, abs(full[k])), (k, run[k], full[k])
    # the engine's configuration, re-evaluated by the oracle from scratch
    pos = eng.positions()
    sd2 = dict(sd, xyz=pos)
    ref = O.Oracle(sd2).energy_full()
    for k in TERMS:
        assert abs(full[k] - ref[k]) <= 1e-10 * max(1e-2, abs(ref[k])), (k, full[k], ref[k])
    # proposals keep bonds within +-10% of r0 (pivot/reptation draw them there;
    # COM and crankshaft preserve them)
    bl = _bond_lengths(pos, sd["mol_start"])
    assert bl.min() >= 0.9 - 1e-9 and bl.max() <= 1.1 + 1e-9


@pytest.mark.parametrize("name,moves", [("small", 1500), ("c0", 1500), ("slab_small", 1500)])
def test_speculative_batches_follow_the_sequential_chain(name, moves):
    """Batches of up to 8 and 16 proposals evaluated together, resolved in
    order with the cross terms of earlier acceptances, make the same Markov
    chain as one proposal at a time (same decisions; positions equal to
    round-off)."""
    sd = systems.config(name)
    kw = dict(move_size=sd["move_size"], seed=23, phantom=sd.get("phantom", 0),
              grafted=sd.get("grafted", 0), bond_len=1.0)
    runs = {}
    for pmax in (1, 8, 16):
        eng = E.Engine(sd)
        eng.set_tuning(pmax=pmax)
        st = eng.mc_run(moves, sd["move_prob"], **kw)
        runs[pmax] = (eng, st)
    e1, s1 = runs[1]
    for pmax in (8, 16):
        ep, sp = runs[pmax]
        print("%s: accepted p1=%s p%d=%s" % (name, list(s1.accepted), pmax, list(sp.accepted)),
              flush=True)
        assert list(s1.attempted) == list(sp.attempted)
        assert list(s1.accepted) == list(sp.accepted)
        assert np.allclose(e1.positions(), ep.positions(), rtol=0, atol=1e-9)
        t1, tp = e1.totals(), ep.totals()
        for k in TERMS:
            assert abs(t1[k] - tp[k]) <= 1e-9 * max(1.0, abs(t1[k])), (k, t1[k], tp[k])
        full = ep.energy_full()
        for k in TERMS:
            if k == "dipole":
                continue
            assert abs(tp[k] - full[k]) <= 1e-9 * max(1.0, abs(full[k])), (k, tp[k], full[k])


@pytest.mark.parametrize("pmax", [1, 16])
def test_queries_shared_by_several_waves_follow_the_same_chain(pmax):
    """Real-space queries spread over several waves (a small split_div: one
    wave per share, four times as many in a batch of fewer than 16
    proposals) make the same chain as one wave per query: same decisions,
    positions equal to round-off, running totals equal to a from-scratch
    evaluation."""
    sd = systems.config("c0")
    kw = dict(move_size=sd["move_size"], seed=31, bond_len=1.0)
    runs = []
    for div in (1536, 64):
        eng = E.Engine(sd)
        eng.set_tuning(pmax=pmax, split_div=div)
        st = eng.mc_run(1500, sd["move_prob"], **kw)
        runs.append((eng, st))
    (e1, s1), (e2, s2) = runs
    print("pmax %d: accepted %s / %s" % (pmax, list(s1.accepted), list(s2.accepted)), flush=True)
    assert list(s1.attempted) == list(s2.attempted)
    assert list(s1.accepted) == list(s2.accepted)
    assert np.allclose(e1.positions(), e2.positions(), rtol=0, atol=1e-9)
    run, full = e2.totals(), e2.energy_full()
    for k in TERMS:
        if k == "dipole":
            continue
        assert abs(run[k] - full[k]) <= 1e-9 * max(1.0, abs(full[k])), (k, run[k], full[k])


def test_device_loop_is_deterministic():
    sd = systems.config("small")
    out = []
    for _ in range(2):
        eng = E.Engine(sd)
        eng.mc_run(500, sd["move_prob"], seed=5)
        out.append(eng.positions())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("name", ["c0", "c1"])
def test_fp32_pair_path_error_per_term(name):
    """FP32 pair math with FP64 accumulation, per-term dE error against the
    FP64 path on identical moves (printed; asserted against 1e-6 relative of
    the per-term scale used by the FP64 parity tests)."""
    sd = systems.config(name)
    e64 = E.Engine(sd, precision=E.PLUM_FP64)
    e32 = E.Engine(sd, precision=E.PLUM_FP32_PAIR)
    full = e64.energy_full()
    rng = np.random.default_rng(7)
    ms = sd["mol_start"]
    chains = [m for m in range(len(ms) - 1) if ms[m + 1] - ms[m] > 1]
    worst = {k: 0.0 for k in TERMS}
    for step in range(40):
        mol = chains[rng.integers(len(chains))]
        a, b = ms[mol], ms[mol + 1]
        idx = np.arange(a, b, dtype=np.int32)
        tr = e64.positions()[a:b] + rng.normal(0, 0.05, size=(b - a, 3))
        t64, d64 = e64.delta(mol, idx, tr)
        t32, d32 = e32.delta(mol, idx, tr)
        for k in TERMS:
            den = max(abs(d64[k]), 1e-3 * abs(full[k]), 1e-2)
            worst[k] = max(worst[k], abs(d32[k] - d64[k]) / den)
        acc = t64 < 1e8 and rng.random() < np.exp(-t64)
        e64.finalize(acc)
        e32.finalize(acc)
    print("fp32[%s] worst rel per term: %s" % (name, " ".join("%s=%.1e" % kv for kv in worst.items())),
          flush=True)
    for k in ("lj", "real", "kspace", "bond", "wall"):
        assert worst[k] <= 1e-6, (k, worst[k])
