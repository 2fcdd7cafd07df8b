"""Configurational-bias trial energies on the GPU (plum_cbmc_trials) against
the oracle's restatement of ForceField::BeadsEnergy (cbmc.cc:5-151), one
trial at a time: insertion steps with and without a grown chain, charged and
neutral chains, a deletion step that leaves a chain and its counterions out
of the system, walls, forbidden trials, and the FP32-pair path."""
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


def test_fp32_pair_path():
    sd = systems.config("small")
    _compare(sd, 16, 4, True, seed=21, precision=E.PLUM_FP32_PAIR, tol=1e-6)


def test_grand_canonical_workload_step():
    """One growth step of the 32-bead chain exchange at its full size (50
    chains + 1600 counterions in a box of 60)."""
    sd = systems.config("c4")
    _compare(sd, 16, 5, True, seed=2)
