"""Parity of the timed device loop, proposal by proposal.

plum_mc_run (k_propose_spec, k_plan_spec, k_compute_spec, k_reduce_spec,
k_resolve_spec) records every proposal it decides: the moved beads, their
trial coordinates, the per-term dE the Metropolis test was given (with the
exact cross terms of the proposals accepted before it in the batch) and the
decision.  Each record is replayed through the CPU oracle, in order, on the
oracle's own copy of the system: the oracle sets the same trial coordinates,
evaluates ForceField::EnergyDifference (force_field.cc:407-434) with the
reference's pairwise sums, and commits or rolls back as the device decided.
Accepted and rejected proposals are checked alike.

Tolerance: the FP64 path agrees per term to 1e-10 relative to max(|dE_ref|,
0.01 kT).  Below 0.01 kT a term changes an acceptance probability
exp(-beta dE) by less than 1%, and an absolute error of 1e-12 kT there moves
it by 1e-12; the strict relative error |got - ref| / |ref| is printed too.
The FP32-pair path agrees to 1e-6 relative to max(|dE_ref|, 1 kT): its pair
terms carry ~1e-7 of relative rounding each, so over the ~10^3 pair images
of a move the absolute error is ~1e-7..1e-6 kT whatever the size of dE, and
1 kT is the scale on which the Metropolis test exp(-beta dE) reads it.
A term that is NaN on both sides (the reference's FENE tether evaluated
beyond its range, log of a negative number) matches; such a move is
rejected on both sides.
"""
import numpy as np
import pytest

from oracle import oracle as O
from plum_amd import engine as E
from plum_amd import systems

pytestmark = pytest.mark.gpu
TERMS = ("lj", "real", "kspace", "self", "dipole", "bond", "wall")
FLOOR = 1e-2  # kT
BIG = 1e8


def _kw(sd, seed):
    return dict(move_size=sd["move_size"], seed=seed, phantom=sd.get("phantom", 0),
                grafted=sd.get("grafted", 0), bond_len=1.0)


def record_run(sd, moves, pmax, seed, precision=E.PLUM_FP64, cap=None):
    eng = E.Engine(sd, precision=precision)
    eng.set_tuning(pmax=pmax)
    cap = moves + 64 if cap is None else cap
    eng.mc_record(cap)
    st = eng.mc_run(moves, sd["move_prob"], **_kw(sd, seed))
    recs, n_dec = eng.mc_records(cap)
    return eng, st, recs, n_dec


def _err(got, ref):
    if np.isnan(got) and np.isnan(ref):
        return 0.0
    return abs(got - ref)


def replay(sd, recs, rel, floor=FLOOR):
    """Replays the records through the oracle; returns the oracle, the worst
    error per term (scaled by max(|ref|, floor)) and the worst strict
    relative error per term."""
    orc = O.Oracle(sd)
    worst = {k: 0.0 for k in TERMS}
    strict = {k: 0.0 for k in TERMS}
    worst["total"] = 0.0
    n_cmp = 0
    for i, r in enumerate(recs):
        if not r["evaluated"]:
            continue
        orc.set_trial(r["mol"], r["idx"], r["trial"])
        t_ref, d_ref = orc.delta(r["mol"])
        if np.isnan(t_ref) or np.isnan(r["total"]):
            assert np.isnan(t_ref) and np.isnan(r["total"]), (i, r["total"], t_ref)
            assert not r["accepted"], i
        elif t_ref >= BIG or r["total"] >= BIG:
            assert (t_ref >= BIG) == (r["total"] >= BIG), (i, r["total"], t_ref)
            assert not r["accepted"], i
        else:
            n_cmp += 1
            for k in TERMS:
                e = _err(r["dE"][k], d_ref[k])
                worst[k] = max(worst[k], e / max(abs(d_ref[k]), floor))
                if abs(d_ref[k]) > 0:
                    strict[k] = max(strict[k], e / abs(d_ref[k]))
                assert e <= rel * max(abs(d_ref[k]), floor), (
                    i, r["move"], r["type"], k, r["dE"][k], d_ref[k], r["accepted"])
            e = abs(r["total"] - t_ref)
            worst["total"] = max(worst["total"], e / max(abs(t_ref), floor))
            assert e <= rel * max(abs(t_ref), floor), (i, r["total"], t_ref)
        orc.finalize(r["mol"], r["accepted"])
    return orc, worst, strict, n_cmp


@pytest.mark.parametrize("name,moves,pmax", [
    ("small", 1500, 1), ("small", 1500, 16),
    ("slab_small", 1000, 1), ("slab_small", 1000, 16),
    ("c0", 520, 1), ("c0", 520, 16),
])
def test_every_decided_proposal_matches_the_oracle(name, moves, pmax):
    """FP64: every decided proposal of the timed loop (accepted or rejected)
    has the oracle's per-term dE to 1e-10, and the replayed chain ends in the
    device's configuration, coordinate for coordinate."""
    sd = systems.config(name)
    eng, st, recs, n_dec = record_run(sd, moves, pmax, seed=41)
    print("[%s p%d] %d moves, %d decided, %d recorded, accepted %d" % (
        name, pmax, st.moves, n_dec, len(recs), sum(r["accepted"] for r in recs)), flush=True)
    assert n_dec == len(recs) >= 500
    assert sum(r["accepted"] for r in recs) > 0
    assert sum(not r["accepted"] for r in recs) > 0
    orc, worst, strict, n_cmp = replay(sd, recs, 1e-10)
    print("[%s p%d] compared %d; worst scaled %s; worst strict %s" % (
        name, pmax, n_cmp, " ".join("%s=%.1e" % kv for kv in worst.items()),
        " ".join("%s=%.1e" % kv for kv in strict.items())), flush=True)
    assert n_cmp >= 450
    assert np.array_equal(orc.positions(), eng.positions())
    # the device's running totals and the oracle's agree too
    run, ref = eng.totals(), orc.running()
    for k in TERMS:
        assert abs(run[k] - ref[k]) <= 1e-9 * max(1.0, abs(ref[k])), (k, run[k], ref[k])


@pytest.mark.parametrize("name", ["small", "c0"])
def test_one_and_sixteen_proposals_record_the_same_decisions(name):
    sd = systems.config(name)
    _, _, r1, _ = record_run(sd, 400, 1, seed=43)
    _, _, r16, _ = record_run(sd, 400, 16, seed=43)
    assert len(r1) == len(r16)
    for a, b in zip(r1, r16):
        assert (a["move"], a["mol"], a["accepted"]) == (b["move"], b["mol"], b["accepted"])
        assert np.array_equal(a["trial"], b["trial"])
        for k in TERMS:
            assert abs(a["dE"][k] - b["dE"][k]) <= 1e-10 * max(abs(a["dE"][k]), FLOOR), k


@pytest.mark.parametrize("name,moves", [("small", 1000), ("c0", 520)])
def test_fp32_pair_path_every_decision_within_1e6(name, moves):
    """The FP32-pair spec kernel (k_compute_spec<1>) on its own chain: every
    decided proposal's per-term dE within 1e-6 kT (1e-6 relative above 1 kT)
    of the oracle's FP64 value."""
    sd = systems.config(name)
    eng, st, recs, n_dec = record_run(sd, moves, 16, seed=47, precision=E.PLUM_FP32_PAIR)
    assert n_dec == len(recs) >= 500
    orc, worst, strict, n_cmp = replay(sd, recs, 1e-6, floor=1.0)
    print("[fp32 %s] compared %d; worst error / max(|dE|, 1 kT): %s; worst strict %s" % (
        name, n_cmp, " ".join("%s=%.1e" % kv for kv in worst.items()),
        " ".join("%s=%.1e" % kv for kv in strict.items())), flush=True)
    assert np.array_equal(orc.positions(), eng.positions())


def test_delayed_acceptance_decisions_match_the_oracle():
    """With the surrogate screening proposals, every exactly evaluated
    proposal (second stage, or a training evaluation of a first-stage
    rejection) still carries the oracle's dE; first-stage rejections leave
    the configuration unchanged."""
    sd = systems.config("small")
    eng = E.Engine(sd)
    eng.surrogate_enable(screen=16, fill=2, train_every=8, train_samples=32, ring=256,
                         r_d=3.0, seed=3)
    eng.mc_run(400, sd["move_prob"], **_kw(sd, 51))     # some online training first
    sd2 = dict(sd, xyz=eng.positions())
    eng.mc_record(4000)
    eng.mc_run(3000, sd["move_prob"], **_kw(sd, 52))
    recs, n_dec = eng.mc_records(4000)
    assert n_dec == len(recs)
    ev = [r for r in recs if r["evaluated"]]
    assert any(not r["pass"] for r in recs) and len(ev) >= 300
    assert all(not r["accepted"] for r in recs if not r["pass"])
    orc, worst, strict, n_cmp = replay(sd2, recs, 1e-10)
    print("[da small] %d decided, %d evaluated; worst %s" % (
        len(recs), n_cmp, " ".join("%s=%.1e" % kv for kv in worst.items())), flush=True)
    assert np.array_equal(orc.positions(), eng.positions())


def test_finalize_rollback_then_device_loop_is_ordered():
    """A host-path rollback (its commit kernel queued on the context's
    stream, the call returning without a wait) followed at once by the device
    loop on its own stream: the loop starts from the rolled-back state."""
    sd = systems.config("c0")
    eng = E.Engine(sd)
    kw = _kw(sd, 61)
    eng.mc_run(200, sd["move_prob"], **kw)
    ms = sd["mol_start"]
    for rep in range(5):
        mol = 2 + rep
        a, b = ms[mol], ms[mol + 1]
        idx = np.arange(a, b, dtype=np.int32)
        eng.delta(mol, idx, eng.positions()[a:b] + 0.3)
        eng.finalize(False)
        st = eng.mc_run(300, sd["move_prob"], **dict(kw, seed=62 + rep))
        assert st.moves == 300 and sum(st.attempted) == 300
    run, full = eng.totals(), eng.energy_full()
    for k in TERMS:
        if k == "dipole":
            continue
        assert abs(run[k] - full[k]) <= 1e-9 * max(1.0, abs(full[k])), (k, run[k], full[k])
