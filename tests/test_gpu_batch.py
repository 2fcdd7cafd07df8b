"""Batched replicas on the GPU: the batched kernels (one launch carries one
trial move of every replica, sweeps replayed from a HIP graph) must produce
exactly the trajectories of the single-system device loop, and each replica's
state must stay consistent with its from-scratch energy."""
import numpy as np
import pytest

from plum_amd import engine as E
from plum_amd import systems

pytestmark = pytest.mark.gpu

R_SMALL = 6
N_MAX = 700


def _mc(sd, seed):
    return dict(move_prob=sd["move_prob"], move_size=sd["move_size"], seed=seed,
                phantom=sd.get("phantom", 0), grafted=sd.get("grafted", 0), bond_len=sd["r0"])


def _replicas(n=R_SMALL, seed=1002):
    return systems.replica_set(n, seed=seed, n_max=N_MAX)


@pytest.mark.parametrize("precision,use_graph", [(E.PLUM_FP64, True), (E.PLUM_FP64, False),
                                                 (E.PLUM_FP32_PAIR, True)])
def test_batch_reproduces_single_system_runs(precision, use_graph):
    sds = _replicas()
    moves = 300   # crosses the 256-move compaction boundary
    singles = [E.Engine(sd, precision=precision) for sd in sds]
    ref_stats = [e.mc_run(moves, seed=11 + r, **{k: v for k, v in _mc(sds[r], 0).items()
                                                  if k != "seed"})
                 for r, e in enumerate(singles)]
    batched = [E.Engine(sd, precision=precision) for sd in sds]
    b = E.Batch(batched)
    st, per = b.mc_run(moves, [_mc(sd, 11 + r) for r, sd in enumerate(sds)],
                       use_graph=use_graph, per_replica=True)
    print("batch", precision, use_graph, "moves", st.moves, "accepted", list(st.accepted))
    assert st.moves == moves * len(sds)
    assert st.overflow_replicas == 0
    for r in range(len(sds)):
        assert list(per[r].attempted) == list(ref_stats[r].attempted), r
        assert list(per[r].accepted) == list(ref_stats[r].accepted), r
        assert np.array_equal(batched[r].positions(), singles[r].positions()), r
        tb, ts = batched[r].totals(), singles[r].totals()
        for t in E.TERMS:
            assert tb[t] == ts[t], (r, t, tb[t], ts[t])
    b.close()
    for e in singles + batched:
        e.close()


def test_batch_running_totals_match_scratch():
    sds = _replicas(8, seed=77)
    engs = [E.Engine(sd) for sd in sds]
    b = E.Batch(engs)
    st = b.mc_run(600, [_mc(sd, 5 + r) for r, sd in enumerate(sds)])
    assert st.moves == 600 * len(sds)
    for r, e in enumerate(engs):
        run = e.totals()
        full = e.energy_full()
        for t in E.TERMS:
            den = max(abs(full[t]), 1e-3 * sum(abs(v) for v in full.values()), 1e-2)
            assert abs(run[t] - full[t]) / den < 1e-9, (r, t, run[t], full[t])
    b.close()
    for e in engs:
        e.close()


def test_batch_autotuner_picks_a_candidate_and_stays_exact():
    sds = _replicas(12, seed=5)
    engs = [E.Engine(sd) for sd in sds]
    b = E.Batch(engs)
    mc = [_mc(sd, 100 + r) for r, sd in enumerate(sds)]
    t = b.tune(mc, moves_per_trial=8)
    print("tuned", t.cell_div, t.blocks_per_cu, t.split_div, t.pmax, "%.1f us/sweep" % t.sweep_us)
    assert t.cell_div in (1, 2, 3)
    assert all(e.info().n_cells[0] == engs[0].info().n_cells[0] for e in engs)
    assert t.blocks_per_cu in (4, 8, 16)
    assert t.split_div in (192, 384, 768)
    assert t.pmax in (1, 2, 4, 8, 16)
    assert t.sweep_us > 0
    b.mc_run(256, mc)
    for r, e in enumerate(engs):
        run, full = e.totals(), e.energy_full()
        for k in E.TERMS:
            den = max(abs(full[k]), 1e-3 * sum(abs(v) for v in full.values()), 1e-2)
            assert abs(run[k] - full[k]) / den < 1e-9, (r, k, run[k], full[k])
    b.close()
    for e in engs:
        e.close()


def test_batch_final_state_matches_oracle():
    from oracle import oracle as O
    sds = _replicas(3, seed=9)
    engs = [E.Engine(sd) for sd in sds]
    b = E.Batch(engs)
    b.mc_run(200, [_mc(sd, 3 + r) for r, sd in enumerate(sds)])
    for r, e in enumerate(engs):
        sd = dict(sds[r])
        sd["xyz"] = e.positions()
        ref = O.Oracle(sd).energy_full()
        got = e.energy_full()
        for t in E.TERMS:
            den = max(abs(ref[t]), 1e-3 * sum(abs(v) for v in ref.values()), 1e-2)
            assert abs(got[t] - ref[t]) / den < 1e-10, (r, t, got[t], ref[t])
    b.close()
    for e in engs:
        e.close()


def test_batch_fp32_pair_drift_stays_within_the_stated_bound():
    """FP32-pair running totals drift by the per-move pair error (<= 1e-6 of a
    term's dE) summed over the accepted moves; the from-scratch resync is what
    removes it."""
    sds = _replicas(6, seed=21)
    engs = [E.Engine(sd, precision=E.PLUM_FP32_PAIR) for sd in sds]
    b = E.Batch(engs)
    moves = 400
    st = b.mc_run(moves, [_mc(sd, 40 + r) for r, sd in enumerate(sds)])
    for r, e in enumerate(engs):
        run, full = e.totals(), e.energy_full()
        for k in ("lj", "real"):
            den = max(abs(full[k]), 1e-2)
            print(r, k, run[k], full[k], abs(run[k] - full[k]) / den)
            assert abs(run[k] - full[k]) / den < 1e-6 * moves, (r, k, run[k], full[k])
        for k in ("kspace", "bond", "self", "dipole", "wall"):
            den = max(abs(full[k]), 1e-3 * sum(abs(v) for v in full.values()), 1e-2)
            assert abs(run[k] - full[k]) / den < 1e-9, (r, k, run[k], full[k])
    b.close()
    for e in engs:
        e.close()


@pytest.mark.parametrize("div", [1, 3])
def test_cell_grid_change_keeps_every_decision_exact(div):
    """Another cell grid (cells of r_s / div) rebuilds the cell list from the
    current positions; the device loop on it still gives every decided
    proposal the oracle's dE to 1e-10 and ends in the oracle's configuration."""
    import test_gpu_record as TR
    sd = systems.config("small")
    eng = E.Engine(sd)
    before = list(eng.info().n_cells)
    eng.set_tuning(cell_div=div)
    after = list(eng.info().n_cells)
    assert (after != before) == (div != 2), (before, after)
    eng.mc_record(1600)
    st = eng.mc_run(1500, sd["move_prob"], **TR._kw(sd, 41))
    recs, n_dec = eng.mc_records(1600)
    assert n_dec == len(recs) >= 500
    orc, worst, strict, n_cmp = TR.replay(sd, recs, 1e-10)
    print("[cell_div %d] cells %s, compared %d, worst %s" % (
        div, after, n_cmp, " ".join("%s=%.1e" % kv for kv in worst.items())), flush=True)
    assert np.array_equal(orc.positions(), eng.positions())
    eng.close()


def test_regrid_through_a_batch_reaches_the_engines_own_loop():
    """An engine whose grid a replica batch changed (after its own device
    loop had uploaded the old grid) runs its own loop on the new grid: the
    running totals still equal the from-scratch energy."""
    sds = _replicas(2, seed=21)
    engs = [E.Engine(sd) for sd in sds]
    engs[0].mc_run(300, **_mc(sds[0], 1))
    b = E.Batch(engs)
    b.set_tuning(cell_div=3)
    b.mc_run(100, [_mc(sd, 7 + r) for r, sd in enumerate(sds)])
    engs[0].mc_run(300, **_mc(sds[0], 2))
    for e in engs:
        run, full = e.totals(), e.energy_full()
        for k in E.TERMS:
            den = max(abs(full[k]), 1e-3 * sum(abs(v) for v in full.values()), 1e-2)
            assert abs(run[k] - full[k]) / den < 1e-9, (k, run[k], full[k])
    b.close()
    for e in engs:
        e.close()
