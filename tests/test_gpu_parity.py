"""GPU parity of the FP64 path against the CPU oracle, through the C ABI.

Tolerance (FP64 path): every per-term total and per-move dE agrees with the
oracle to 1e-10 relative.  Where a dE term is a small difference of large
sums, the relative error is taken against the larger of |ref| and 1e-3 of
the term's total magnitude (the rounding of either side scales with the sum,
not with the difference), and never against less than 1e-2 kT (terms that
are zero up to rounding, e.g. springs built at exactly r0, carry ~1e-28 of
noise on both sides).
"""
import numpy as np
import pytest

from oracle import oracle as O
from plum_amd import engine as E
from plum_amd import systems

pytestmark = pytest.mark.gpu
TERMS = ("lj", "real", "kspace", "self", "dipole", "bond", "wall")
REL = 1e-10


FLOOR = 1e-2  # kT


def _den(ref, scale=0.0):
    return max(abs(ref), 1e-3 * abs(scale), FLOOR)


def _close(a, b, scale=0.0, rel=REL):
    return abs(a - b) <= rel * _den(b, scale)


def _report(tag, got, ref, scale=None):
    errs = {}
    for k in TERMS:
        s = abs(scale[k]) if scale else 0.0
        errs[k] = abs(got[k] - ref[k]) / _den(ref[k], s)
    print(tag, " ".join("%s=%.1e" % (k, v) for k, v in errs.items()), flush=True)
    return errs


@pytest.mark.parametrize("name", ["tiny", "small", "slab_small", "c0"])
def test_full_energy_matches_oracle(name):
    sd = systems.config(name)
    eng = E.Engine(sd)
    orc = O.Oracle(sd)
    got, ref = eng.energy_full(), orc.energy_full()
    info, oinfo = eng.info(), orc.ewald_info()
    assert info.real_cutoff == oinfo.real_cutoff
    assert list(info.repl_cell) == list(oinfo.repl_cell)
    assert info.n_k_full == oinfo.n_k_full
    _report("full[%s]" % name, got, ref)
    for k in TERMS:
        assert _close(got[k], ref[k]), (k, got[k], ref[k])
    # running totals start at the from-scratch energy
    run = eng.totals()
    for k in TERMS:
        assert _close(run[k], ref[k]), k


def _replay(name, n_moves, seed, move_size=0.5, kinds=None):
    """Drive both sides with the reference move generators (oracle mt19937);
    accept with the oracle's dE; compare per-move dE term by term."""
    sd = systems.config(name)
    eng = E.Engine(sd)
    orc = O.Oracle(sd)
    scale = orc.energy_full()
    rng = O.MT19937(seed)
    ms = sd["mol_start"]
    chains = [m for m in range(len(ms) - 1) if ms[m + 1] - ms[m] > 1]
    ions = [m for m in range(len(ms) - 1) if ms[m + 1] - ms[m] == 1
            and m >= sd.get("phantom", 0)]
    kinds = kinds or (["com", "pivot", "crankshaft", "reptation", "bead"] if sd["npbc"] == 3
                      else ["com", "pivot", "crankshaft"])
    worst = {k: 0.0 for k in TERMS}
    worst_tot = 0.0
    n_acc = 0
    for step in range(n_moves):
        kind = kinds[step % len(kinds)]
        if kind == "bead":
            mol = ions[rng.next() % len(ions)]
        else:
            mol = chains[rng.next() % len(chains)]
        idx, tr = O.propose(orc, rng, kind, mol, move_size=move_size, bond_len=1.0)
        ref_tot, ref = orc.delta(mol)
        got_tot, got = eng.delta(mol, idx, tr)
        for k in TERMS:
            worst[k] = max(worst[k], abs(got[k] - ref[k]) / _den(ref[k], scale[k]))
            assert _close(got[k], ref[k], scale[k]), (step, kind, k, got[k], ref[k])
        den = _den(ref_tot, sum(abs(v) for v in scale.values()))
        worst_tot = max(worst_tot, abs(got_tot - ref_tot) / den)
        acc = ref_tot < 1e8 and (rng.next() / 4294967295.0) < np.exp(-ref_tot)
        n_acc += acc
        orc.finalize(mol, acc)
        eng.finalize(acc)
    print("replay[%s] moves=%d accepted=%d worst_rel: %s total=%.1e" % (
        name, n_moves, n_acc, " ".join("%s=%.1e" % kv for kv in worst.items()), worst_tot),
        flush=True)
    return sd, eng, orc, n_acc


@pytest.mark.parametrize("name,n", [("tiny", 60), ("small", 60), ("slab_small", 60), ("c0", 30)])
def test_per_move_delta_matches_oracle(name, n):
    sd, eng, orc, n_acc = _replay(name, n, seed=17)
    assert n_acc > 0
    # positions stay identical (both sides commit the same trial coordinates)
    assert np.array_equal(eng.positions(), orc.positions())
    # running totals agree with the oracle's incremental totals
    run, orun = eng.totals(), orc.running()
    full = orc.energy_full()
    for k in TERMS:
        assert _close(run[k], orun[k], full[k]), (k, run[k], orun[k])
    # and the engine's from-scratch energy agrees with the oracle's
    got = eng.energy_full()
    for k in TERMS:
        assert _close(got[k], full[k]), (k, got[k], full[k])


def test_resync_drift_is_rounding_only():
    sd, eng, orc, _ = _replay("small", 80, seed=5)
    drift = eng.resync()
    full = orc.energy_full()
    for k in TERMS:
        if k == "dipole":
            continue
        assert abs(drift[k]) <= 1e-11 * max(1.0, abs(full[k])), (k, drift[k])


def test_overlap_and_wall_sentinels():
    sd = systems.config("slab_small")
    eng = E.Engine(sd)
    ms = sd["mol_start"]
    # push the first chain's second bead through the z = 0 wall: forbidden
    mol = 0
    idx = np.array([ms[mol] + 1], dtype=np.int32)
    tr = sd["xyz"][ms[mol] + 1].copy()
    tr[2] = -0.5
    tot, t = eng.delta(mol, idx, tr[None, :])
    assert tot >= 1e8 and t["wall"] == 1e8
    eng.finalize(False)
    # exact overlap with another bead: LJ sentinel, early exit
    tr2 = sd["xyz"][ms[1]].copy()
    tot, t = eng.delta(mol, idx, tr2[None, :])
    assert tot >= 1e8 and t["lj"] >= 1e8 and t["real"] == 0 and t["kspace"] == 0
    eng.finalize(False)


def test_call_order_errors():
    sd = systems.config("tiny")
    eng = E.Engine(sd)
    with pytest.raises(E.PlumError):
        eng.finalize(True)            # no pending move
    with pytest.raises(E.PlumError):
        eng.delta(99, [0], np.zeros((1, 3)))   # molecule out of range
    ms = sd["mol_start"]
    with pytest.raises(E.PlumError):
        eng.delta(0, [ms[1]], np.zeros((1, 3)))  # bead not in molecule


def test_large_system_properties():
    """Large bulk system (40,000 particles): properties that hold at any size.
    dE of an accepted move equals the change of the from-scratch energy, and
    moving back gives exactly the opposite dE."""
    sd = systems.config("c1")
    eng = E.Engine(sd)
    info = eng.info()
    print("c1 info: rc=%g nk_half=%d cells=%s cap=%d" % (
        info.real_cutoff, info.n_k_half, list(info.n_cells), info.cell_cap), flush=True)
    assert info.real_cutoff == 12.0
    e0 = eng.energy_full()
    ms = sd["mol_start"]
    rng = np.random.default_rng(3)
    mol = 17
    a, b = ms[mol], ms[mol + 1]
    idx = np.arange(a, b, dtype=np.int32)
    shift = np.array([0.31, -0.22, 0.17])
    tr = sd["xyz"][a:b] + shift
    tot, t = eng.delta(mol, idx, tr)
    eng.finalize(True)
    e1 = eng.energy_full()
    for k in ("lj", "real", "kspace", "bond"):
        assert abs(t[k] - (e1[k] - e0[k])) <= 1e-10 * max(abs(e0[k]) + abs(e1[k]), FLOOR), k
    tot_b, tb = eng.delta(mol, idx, sd["xyz"][a:b])
    for k in ("lj", "real", "kspace", "bond"):
        assert abs(tb[k] + t[k]) <= 1e-10 * _den(t[k], e0[k]), k
    eng.finalize(True)
    drift = eng.resync()
    for k in ("lj", "real", "kspace", "bond"):
        assert abs(drift[k]) <= 1e-11 * max(1.0, abs(e0[k])), (k, drift[k])
    del rng
