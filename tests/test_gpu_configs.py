"""The stated workloads at their full size on the GPU: the 40,000-particle
bulk system the throughput is quoted on, the confined brush (64 x 50 grafted
chains, 400 charged surface sites, slab Ewald), all 256 batched replicas, and
the reference's own NVT example decks replayed move for move against the
oracle."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from plum_amd import engine as E
from plum_amd import host as H
from plum_amd import systems

from deck import deck_sysdef

pytestmark = pytest.mark.gpu
TERMS = ("lj", "real", "kspace", "self", "dipole", "bond", "wall")
FLOOR = 1e-2
HERE = os.path.dirname(os.path.abspath(__file__))
EX = os.path.join(HERE, "golden", "examples")


def _kw(sd, seed):
    return dict(move_size=sd["move_size"], seed=seed, phantom=sd.get("phantom", 0),
                grafted=sd.get("grafted", 0), bond_len=sd["r0"])


def _running_is_scratch(eng, skip=("dipole",), rel=1e-9):
    run, full = eng.totals(), eng.energy_full()
    for k in TERMS:
        if k in skip:
            continue
        assert abs(run[k] - full[k]) <= rel * max(1.0, abs(full[k])), (k, run[k], full[k])


def _replay_host_path(sd, recs, rel, precision=E.PLUM_FP64, floor=FLOOR):
    """Replays recorded device-loop decisions through the host path
    (plum_energy_difference: the per-move kernels, oracle-checked move by
    move at 1e-10 on the small systems); returns the worst error scaled by
    max(|dE|, floor)."""
    ref = E.Engine(sd, precision=precision)
    worst = {k: 0.0 for k in TERMS}
    for i, r in enumerate(recs):
        if not r["evaluated"]:
            continue
        t, d = ref.delta(r["mol"], r["idx"], r["trial"])
        if t >= 1e8 or r["total"] >= 1e8:
            assert (t >= 1e8) == (r["total"] >= 1e8), i
        else:
            for k in TERMS:
                e = abs(r["dE"][k] - d[k]) / max(abs(d[k]), floor)
                worst[k] = max(worst[k], e)
                assert e <= rel, (i, r["type"], k, r["dE"][k], d[k])
        ref.finalize(r["accepted"])
    return ref, worst


def test_bulk40k_device_loop_matches_host_path_and_sequential_chain():
    """The timed workload: 2,000 device-loop moves; every decided proposal's
    per-term dE equals the host path's on the same trial coordinates; one
    proposal per batch makes the same decisions as sixteen; running totals
    equal the from-scratch energy."""
    sd = systems.config("c1")
    runs = {}
    for pmax in (16, 1):
        eng = E.Engine(sd)
        eng.set_tuning(pmax=pmax)
        eng.mc_record(2100)
        st = eng.mc_run(2000, sd["move_prob"], **_kw(sd, 71))
        recs, n_dec = eng.mc_records(2100)
        assert n_dec == len(recs) == 2000
        runs[pmax] = (eng, st, recs)
        print("c1 p%d: accepted %s, %.0f moves/s" % (pmax, list(st.accepted),
                                                     st.moves / st.seconds), flush=True)
    e16, s16, r16 = runs[16]
    e1, s1, r1 = runs[1]
    assert [(r["move"], r["accepted"]) for r in r16] == [(r["move"], r["accepted"]) for r in r1]
    assert np.allclose(e16.positions(), e1.positions(), rtol=0, atol=1e-9)
    _running_is_scratch(e16)
    ref, worst = _replay_host_path(sd, r16, 1e-10)
    print("c1 spec vs host path, worst scaled: %s" % " ".join("%s=%.1e" % kv for kv in worst.items()),
          flush=True)
    assert np.array_equal(ref.positions(), e16.positions())


def test_bulk40k_fp32_spec_kernel_error_per_term():
    """FP32-pair spec kernel on the timed workload: each decided proposal's
    per-term dE against the FP64 host path on the same trial coordinates,
    within 1e-6 of max(|dE|, 1 kT) (the absolute FP32 pair error of a move
    is ~1e-7..1e-6 kT; 1 kT is the scale of the Metropolis test)."""
    sd = systems.config("c1")
    eng = E.Engine(sd, precision=E.PLUM_FP32_PAIR)
    eng.mc_record(600)
    eng.mc_run(500, sd["move_prob"], **_kw(sd, 73))
    recs, n_dec = eng.mc_records(600)
    assert len(recs) == 500
    _, worst = _replay_host_path(sd, recs, 1e-6, floor=1.0)
    print("c1 fp32 spec vs fp64 host path, worst scaled: %s" % " ".join(
        "%s=%.1e" % kv for kv in worst.items()), flush=True)


def test_brush_full_size_device_loop():
    """The confined brush at its stated size (3,200 grafted beads, 3,200
    counterions, 400 surface sites, walls, slab Ewald with the dipole lag):
    p1 and p16 decide alike, running totals equal the from-scratch energy,
    and the first decided proposals (pivots of 50 beads included) carry the
    oracle's per-term dE."""
    sd = systems.config("c3")
    runs = {}
    for pmax in (16, 1):
        eng = E.Engine(sd)
        eng.set_tuning(pmax=pmax)
        eng.mc_record(3100)
        st = eng.mc_run(3000, sd["move_prob"], **_kw(sd, 81))
        recs, _ = eng.mc_records(3100)
        runs[pmax] = (eng, recs)
        print("c3 p%d: accepted %s of %s, %.0f moves/s" % (
            pmax, list(st.accepted), list(st.attempted), st.moves / st.seconds), flush=True)
    (e16, r16), (e1, r1) = runs[16], runs[1]
    assert [(r["move"], r["accepted"]) for r in r16] == [(r["move"], r["accepted"]) for r in r1]
    assert np.allclose(e16.positions(), e1.positions(), rtol=0, atol=1e-9)
    _running_is_scratch(e16)
    # the oracle from the initial configuration (pair sums recomputed per
    # move, no full-energy pass: the dipole lag state starts from M_z)
    orc = O.Oracle(sd, init_energy=False)
    for i, r in enumerate(r16[:6]):
        orc.set_trial(r["mol"], r["idx"], r["trial"])
        t, d = orc.delta(r["mol"])
        if t < 1e8 and r["total"] < 1e8:
            for k in TERMS:
                assert abs(r["dE"][k] - d[k]) <= 1e-10 * max(abs(d[k]), FLOOR), (i, k, r["dE"][k], d[k])
        else:
            assert (t >= 1e8) == (r["total"] >= 1e8)
        orc.finalize(r["mol"], r["accepted"])
    orc.close()


def test_all_256_replicas_in_one_launch():
    """All 256 batched replicas (2.6M particles, FP32 pairs): every replica
    makes its moves with no cell overflow left, k space / springs stay exact
    and the FP32 pair drift stays within its stated bound."""
    sds = systems.replica_set(256)
    engs = [E.Engine(sd, precision=E.PLUM_FP32_PAIR) for sd in sds]
    b = E.Batch(engs)
    mcs = [dict(move_prob=sd["move_prob"], move_size=1.0, seed=900 + r, bond_len=1.0)
           for r, sd in enumerate(sds)]
    moves = 64
    st, per = b.mc_run(moves, mcs, per_replica=True)
    print("256 replicas: %d moves in %.2f s, overflow %d" % (st.moves, st.seconds,
                                                             st.overflow_replicas), flush=True)
    assert st.moves == moves * 256
    assert all(p.moves == moves for p in per)
    for r, e in enumerate(engs):
        run, full = e.totals(), e.energy_full()
        for k in ("kspace", "bond", "self", "wall"):
            den = max(abs(full[k]), 1e-3 * sum(abs(v) for v in full.values()), 1e-2)
            assert abs(run[k] - full[k]) / den < 1e-9, (r, k, run[k], full[k])
        for k in ("lj", "real"):
            den = max(abs(full[k]), 1e-2)
            assert abs(run[k] - full[k]) / den < 1e-6 * moves, (r, k, run[k], full[k])
    b.close()
    for e in engs:
        e.close()


@pytest.mark.parametrize("deck,steps", [("bulk_nvt", 400), ("confined_nvt", 400)])
def test_reference_nvt_example_decks_replay_on_the_oracle(deck, steps):
    """The reference's own NVT decks (bulk: 16 rigid chains + ions, WCA,
    l_B 2.5, alpha 0.001 so every pair sums 125 images; confined: 18 charged
    surface sites, walls, slab dipole): the host driver with GPU energies and
    the oracle take the same decisions from the same seed and end in the
    same configuration with the same energy."""
    path = os.path.join(EX, deck, "run.in")
    sd, mv = deck_sysdef(path)
    sim = H.HostSimulation(path, 2017)
    att, acc = sim.run(steps)
    orc = O.Oracle(sd)
    rng = O.MT19937(2017)
    oatt, oacc = [0] * 5, [0] * 5
    for _ in range(steps):
        inf = O.mc_step(orc, rng, sd["move_prob"], sd["move_size"], mv["phantom"], mv["grafted"],
                        mv["bond_len"], mv["vary_bond"])
        if inf.attempted:
            oatt[inf.move_type] += 1
            oacc[inf.move_type] += inf.accepted
    print(deck, "host", att, acc, "oracle", oatt, oacc, flush=True)
    assert att == oatt and acc == oacc and sum(acc) > 0
    pos = sim.positions(len(sd["q"]))
    assert np.max(np.abs(pos - orc.positions())) <= 1e-9
    run, full = sim.energies()
    ref = orc.energy_full()
    for k in ref:
        assert abs(full[k] - ref[k]) <= 1e-10 * max(abs(ref[k]), 1e-2), (k, full[k], ref[k])
    orc.close()
    sim.close()
