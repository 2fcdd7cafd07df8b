"""Configurational-bias trial energies on the GPU (plum_cbmc_trials) against
the oracle's restatement of ForceField::BeadsEnergy (cbmc.cc:5-151), one
trial at a time: insertion steps with and without a grown chain, charged and
neutral chains, a deletion step that leaves a chain and its counterions out
of the system, walls, forbidden trials, the slab dipole term, and the
FP32-pair path."""
import numpy as np
import pytest

from oracle import oracle as O
from plum_amd import engine as E
from plum_amd import systems

pytestmark = pytest.mark.gpu


def _walk(rng, start, n, r0=1.0):
    out = [np.asarray(start, dtype=float)]
    while len(out) < n:
        v = rng.normal(size=3)
        out.append(out[-1] + r0 * v / np.linalg.norm(v))
    return np.array(out)


def _trials(rng, end, K, box, charged):
    v = rng.normal(size=(K, 3))
    t = end + 0.97 * v / np.linalg.norm(v, axis=1)[:, None]
    ions = rng.uniform(0, 1, (K, 3)) * box if charged else None
    return t, ions


def _compare(sd, K, cur, charged, seed, exclude_mols=None, precision=None, tol=1e-10,
             plant_overlap=False):
    rng = np.random.default_rng(seed)
    box = np.asarray(sd["box"], dtype=float)
    q_c, q_i = (-1.0, 1.0) if charged else (0.0, 0.0)
    if sd["npbc"] == 2:
        start = np.r_[rng.uniform(0, 1, 2) * box[:2], box[2] / 2]
    else:
        start = rng.uniform(0, 1, 3) * box
    grown = _walk(rng, start, max(cur, 1))[:cur]
    gi = rng.uniform(0.1, 0.9, (cur, 3)) * box if charged else None
    end = grown[-1] if cur else start
    tri, ions = _trials(rng, end, K, box, charged)
    if cur == 0:
        tri = rng.uniform(0.05, 0.95, (K, 3)) * box
    if plant_overlap:
        tri[1] = np.asarray(sd["xyz"][3]) + 1e-3
    ex_beads = None
    if exclude_mols is not None:
        ms = sd["mol_start"]
        ex_beads = (int(ms[exclude_mols[0]]), int(ms[exclude_mols[1] + 1]) - 1)
    kw = {} if precision is None else {"precision": precision}
    eng = E.Engine(sd, **kw)
    got = eng.cbmc_trials(tri, 0, q_c, grown=grown if cur else None, ions=ions, type_ion=1,
                          q_ion=q_i, grown_ions=gi, exclude=ex_beads)
    eng.close()
    o = O.Oracle(sd, init_energy=False)     # the bead energy needs no running totals
    want = np.array([o.cbmc_beads_energy(tri[t], 0, q_c, grown=grown if cur else None,
                                         b2=None if ions is None else ions[t], t2=1, q2=q_i,
                                         grown_ions=gi, exclude_mols=exclude_mols)
                     for t in range(K)])
    o.close()
    err = np.abs(got - want) / np.maximum(1.0, np.abs(want))
    print("K=%d cur=%d charged=%d: energies %s, worst rel err %.2e" % (
        K, cur, charged, np.array2string(want[:4], precision=4), err.max()), flush=True)
    assert err.max() <= tol, (got, want)
    return got, want


@pytest.mark.parametrize("cur,charged", [(0, True), (1, True), (4, True), (3, False)])
def test_insertion_step_matches_oracle(cur, charged):
    _compare(systems.config("small"), 16, cur, charged, seed=cur + 10 * charged)


def test_deletion_step_leaves_the_chain_and_counterions_out():
    sd = systems.config("gc_small")
    L = sd["gc"]["chain_len"]
    _compare(sd, 8, 3, True, seed=4, exclude_mols=(1 * (1 + L), 1 * (1 + L) + L))


def test_walls_and_forbidden_trials():
    sd = systems.config("slab_small", dipole=0)
    got, want = _compare(sd, 12, 3, True, seed=8, plant_overlap=True)
    assert got[1] == 1e8 and want[1] == 1e8


@pytest.mark.parametrize("cur,charged,exclude", [(0, True, None), (3, True, None), (2, False, None),
                                                 (4, True, (5, 9))])
def test_slab_dipole_term_matches_oracle(cur, charged, exclude):
    """Slab Ewald with the dipole correction: each trial's energy carries
    DipoleEDiff (potential_ewald_coul.cc:473-530) over the system less the
    excluded molecules, the grown chain and its counterions."""
    sd = systems.config("slab_small")
    assert sd["dipole"]
    _compare(sd, 12, cur, charged, seed=30 + cur, exclude_mols=exclude)


def test_fp32_pair_path():
    sd = systems.config("small")
    _compare(sd, 16, 4, True, seed=21, precision=E.PLUM_FP32_PAIR, tol=1e-6)


def test_grand_canonical_workload_step():
    """One growth step of the 32-bead chain exchange at its full size (50
    chains + 1600 counterions in a box of 60)."""
    sd = systems.config("c4")
    _compare(sd, 16, 5, True, seed=2)


def _stepwise(eng, mode, L, K, first, disp, ions, sel, actual, charged, exclude, beta=1.0):
    """The growth loop of the reference with each step's trial energies from
    plum_cbmc_trials: the weight, the step where it stopped, the chain."""
    q_c, q_i = (-1.0, 1.0) if charged else (0.0, 0.0)
    chain = np.zeros((L, 3))
    cions = np.zeros((L, 3))
    if mode == 1:
        chain[0], cions[0] = actual[0], actual[L]
    else:
        chain[0], cions[0] = first[0], first[1]
    e0 = eng.cbmc_trials(chain[:1], 0, q_c, ions=cions[:1] if charged else None, type_ion=0,
                         q_ion=q_i, exclude=exclude)[0]
    W = np.exp(-beta * e0)
    if mode == 0 and W <= 0:
        return W, 0, chain, cions
    for s in range(1, L):
        tri = chain[s - 1] + disp[s]
        ion = ions[s].copy()
        if mode == 1:
            tri[0], ion[0] = actual[s], actual[L + s]
        e = eng.cbmc_trials(tri, 0, q_c, grown=chain[:s], ions=ion if charged else None,
                            type_ion=0, q_ion=q_i, grown_ions=cions[:s] if charged else None,
                            exclude=exclude)
        w = np.exp(-beta * e)
        Wi = 0.0
        for x in w:
            Wi += x
        W *= Wi / K
        if mode != 1 and W <= 0:
            return W, s, chain, cions
        cb = 0
        if mode != 1:
            rn, cum = sel[s] * Wi, w[0]
            while cum < rn and cb < K - 1:
                cb += 1
                cum += w[cb]
        chain[s], cions[s] = tri[cb], ion[cb]
    return W, -1, chain, cions


@pytest.mark.parametrize("name,mode,charged", [("small", 0, True), ("small", 0, False),
                                               ("gc_small", 1, True), ("small", 2, True),
                                               ("c4", 0, True), ("c4", 1, True),
                                               ("slab_small", 0, True), ("slab_small", 2, True)])
def test_whole_growth_matches_the_stepwise_loop(name, mode, charged):
    """plum_cbmc_grow (every step on the device) against the reference's
    growth loop driven step by step through plum_cbmc_trials on the same
    draws: the same Rosenbluth weight, the same chosen beads."""
    sd = systems.config(name)
    rng = np.random.default_rng(40 + mode)
    L = sd["gc"]["chain_len"] if "gc" in sd else 6
    K = 16 if name == "c4" else 8
    box = np.asarray(sd["box"])
    first = rng.uniform(0.1, 0.9, (2, 3)) * box
    v = rng.normal(size=(L, K, 3))
    disp = 0.98 * v / np.linalg.norm(v, axis=2)[..., None]
    ions = rng.uniform(0, 1, (L, K, 3)) * box
    sel = rng.uniform(0, 1, L)
    actual, exclude = None, None
    if mode == 1:
        # the second exchanged chain and its counterions (molecules 1 + L .. 2 L + 1)
        ms = sd["mol_start"]
        m = 1 + L
        actual = np.asarray(sd["xyz"])[ms[m]:ms[m + 1 + L]]
        exclude = (int(ms[m]), int(ms[m + 1 + L]) - 1)
    eng = E.Engine(sd)
    W, stop, chain = eng.cbmc_grow(mode, L, K, 0, -1.0 if charged else 0.0, first=first,
                                   disp=disp, ions=ions, sel=sel, actual=actual, type_ion=0,
                                   q_ion=1.0, charged=charged, exclude=exclude)
    W2, stop2, ch2, ci2 = _stepwise(eng, mode, L, K, first, disp, ions, sel, actual, charged,
                                    exclude)
    eng.close()
    print("%s mode %d: W %.6e vs %.6e, stop %d vs %d" % (name, mode, W, W2, stop, stop2),
          flush=True)
    assert stop == stop2
    assert abs(W - W2) <= 1e-9 * abs(W2) or (W == 0 and W2 == 0)
    if stop < 0:
        assert np.max(np.abs(chain[:L] - ch2)) <= 1e-12
        if charged:
            assert np.max(np.abs(chain[L:] - ci2)) <= 1e-12


def test_growth_stops_where_the_weight_vanishes():
    sd = systems.config("small")
    L, K = 5, 4
    rng = np.random.default_rng(3)
    first = np.array([np.asarray(sd["xyz"][7]) + 1e-4, [1.0, 2.0, 3.0]])  # on top of a bead
    v = rng.normal(size=(L, K, 3))
    disp = v / np.linalg.norm(v, axis=2)[..., None]
    ions = rng.uniform(0, 14, (L, K, 3))
    sel = rng.uniform(0, 1, L)
    eng = E.Engine(sd)
    W, stop, _ = eng.cbmc_grow(0, L, K, 0, -1.0, first=first, disp=disp, ions=ions, sel=sel,
                               q_ion=1.0)
    assert W == 0 and stop == 0                     # insertion: returns after the first bead
    W, stop, _ = eng.cbmc_grow(2, L, K, 0, -1.0, first=first, disp=disp, ions=ions, sel=sel,
                               q_ion=1.0)
    assert W == 0 and stop == 1                     # test insertion: stops after step 1
    eng.close()
