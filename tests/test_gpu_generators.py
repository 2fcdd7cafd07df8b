"""The device loop's move generators (k_propose_spec, restating the generators
of molecule.cc:103-312 on Philox4x32-10) against the rules of the reference's
generators, proposal by proposal.

plum_mc_run records every proposal it decides with its trial coordinates; the
coordinates before each proposal come from replaying the recorded decisions
on the oracle's copy of the system.  The reference draws from mt19937 and the
device from Philox, so the two are compared through what each rule fixes
exactly, and through the distributions of the rules' random parts:
- BeadTranslate (molecule.cc:103-134): a step of exactly 3 move_size along a
  direction uniform on the sphere (randSphere, misc.cc:95-109);
- COMTranslate (:135-152): one rigid displacement per molecule, each component
  uniform on [-move_size/2, move_size/2];
- Pivot (:153-238): the pivot bead stays, every bond is redrawn with a length
  within +-10% of the equilibrium length (vary_bond);
- Crankshaft (:239-266): a rotation about the axis through the first and last
  (unmoved) beads, so each moved bead keeps its distances to both;
- RandomReptation (:267-312): every bead takes its neighbour's slot in one
  direction and the end bead steps by a length uniform within +-10% of the
  equilibrium length along a direction uniform on the sphere.
The statistical bounds are five standard errors.
"""
import numpy as np
import pytest

from oracle import oracle as O
from plum_amd import engine as E
from plum_amd import systems

pytestmark = pytest.mark.gpu

MIX = [0.2, 0.2, 0.2, 0.2, 0.2]   # bead, COM, pivot, crankshaft, reptation
R0 = 1.0
Z = 5.0


def _uniform_sphere(u, label):
    """Unit vectors u (n x 3) drawn uniformly on the sphere: mean 0 and
    <u_k^2> = 1/3 per component (variance of u_k^2 is 4/45)."""
    n = len(u)
    assert n > 200, (label, n)
    assert np.allclose(np.linalg.norm(u, axis=1), 1.0, atol=1e-12), label
    m = u.mean(axis=0)
    assert np.all(np.abs(m) < Z * np.sqrt(1.0 / 3.0 / n)), (label, m)
    m2 = (u * u).mean(axis=0)
    assert np.all(np.abs(m2 - 1.0 / 3.0) < Z * np.sqrt(4.0 / 45.0 / n)), (label, m2)


def test_device_generators_follow_the_reference_rules():
    sd = systems.config("small")
    msz = sd["move_size"]
    ms = np.asarray(sd["mol_start"])
    moves = 4000
    eng = E.Engine(sd)
    cap = moves + 64
    eng.mc_record(cap)
    st = eng.mc_run(moves, MIX, move_size=msz, seed=31, bond_len=R0)
    recs, n_dec = eng.mc_records(cap)
    print("decided", n_dec, "attempted", list(st.attempted), "accepted", list(st.accepted), flush=True)
    assert n_dec == len(recs) and n_dec >= moves

    orc = O.Oracle(sd)
    bead_dir, rep_dir, rep_len, com = [], [], [], []
    seen = np.zeros(5, dtype=int)
    for i, r in enumerate(recs):
        if not r["evaluated"]:
            continue
        idx = np.asarray(r["idx"])
        new = np.asarray(r["trial"], dtype=float).reshape(-1, 3)
        pos = orc.positions()
        old = pos[idx]
        kind = r["type"]
        mol = r["mol"]
        a, b = ms[mol], ms[mol + 1]
        if len(idx) > 0:
            seen[kind] += 1
        if kind == 0:
            assert len(idx) == 1 and b - a == 1, i
            d = new[0] - old[0]
            step = np.linalg.norm(d)
            assert abs(step - 3 * msz) <= 1e-12 * 3 * msz, (i, step)
            bead_dir.append(d / step)
        elif kind == 1:
            assert list(idx) == list(range(a, b)), i
            d = new - old
            assert np.allclose(d, d[0], rtol=0, atol=1e-12), i
            assert np.all(np.abs(d[0]) <= 0.5 * msz * (1 + 1e-12)), (i, d[0])
            com.append(d[0] / (0.5 * msz))
        elif kind == 2:
            assert list(idx) == list(range(a, b)), i
            assert np.any(np.all(new == old, axis=1)), i   # the pivot bead stays
            bl = np.linalg.norm(np.diff(new, axis=0), axis=1)
            assert bl.min() >= 0.9 * R0 - 1e-12 and bl.max() <= 1.1 * R0 + 1e-12, (i, bl.min(), bl.max())
        elif kind == 3:
            if len(idx) == 0:
                continue
            assert np.all(np.diff(idx) == 1), i
            first, last = pos[idx[0] - 1], pos[idx[-1] + 1]
            assert a <= idx[0] - 1 and idx[-1] + 1 < b, i
            for p in (first, last):
                dn = np.linalg.norm(new - p, axis=1)
                do = np.linalg.norm(old - p, axis=1)
                assert np.allclose(dn, do, rtol=0, atol=1e-10), (i, np.abs(dn - do).max())
        elif kind == 4:
            assert list(idx) == list(range(a, b)), i
            fwd = np.array_equal(new[:-1], old[1:])
            bwd = np.array_equal(new[1:], old[:-1])
            assert fwd != bwd, i   # exactly one direction (chains are longer than one bead)
            e = len(idx) - 1 if fwd else 0
            d = new[e] - old[e]
            step = np.linalg.norm(d)
            assert 0.9 * R0 - 1e-12 <= step <= 1.1 * R0 + 1e-12, (i, step)
            rep_dir.append(d / step)
            rep_len.append(step)
        orc.set_trial(mol, r["idx"], r["trial"])
        orc.delta(mol)
        orc.finalize(mol, r["accepted"])
    print("proposals per kind", list(seen), flush=True)
    assert np.all(seen > 300), seen

    _uniform_sphere(np.array(bead_dir), "bead translation")
    _uniform_sphere(np.array(rep_dir), "reptation end")
    c = np.array(com)
    n = len(c)
    assert np.all(np.abs(c.mean(axis=0)) < Z * np.sqrt(1.0 / 3.0 / n)), c.mean(axis=0)
    assert np.all(np.abs((c * c).mean(axis=0) - 1.0 / 3.0) < Z * np.sqrt(4.0 / 45.0 / n)), (c * c).mean(axis=0)
    # the step length of a reptation end: R0 + (R0 / 5) (U - 1/2)
    s = np.array(rep_len)
    assert abs(s.mean() - R0) < Z * (0.2 * R0 / np.sqrt(12.0)) / np.sqrt(len(s)), s.mean()
    assert abs(s.var() - (0.2 * R0) ** 2 / 12.0) < Z * (0.2 * R0) ** 2 * np.sqrt(1.0 / 180.0 / len(s)), s.var()
