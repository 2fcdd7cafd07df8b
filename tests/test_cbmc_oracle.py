"""CPU checks that pin the oracle's configurational-bias bead energy
(orc_cbmc_beads_energy, restating ForceField::BeadsEnergy of cbmc.cc:5-151)
by what the energy function fixes: the energy of placing a trial chain bead
and its counterion next to the chain grown so far equals the difference of
two full system energies (the system with the grown chain, and with the two
new beads added), less the bond and the bonded LJ pair that BeadsEnergy
leaves out; and leaving molecules out by index range equals removing them
from the system."""
import numpy as np
import pytest

from oracle import oracle as O
from plum_amd import systems

NONBOND = ("lj", "real", "kspace", "self", "wall")


def _with_chain(sd, chain, ions, q_c, q_i, t_c=0, t_i=1):
    """sd plus a chain (beads `chain`) and single-bead counterions `ions`."""
    xyz = [sd["xyz"]]
    q = [sd["q"]]
    ty = [sd["type"]]
    ms = list(sd["mol_start"])
    n = ms[-1]
    if len(chain):
        xyz.append(np.asarray(chain).reshape(-1, 3))
        q.append(np.full(len(chain), q_c))
        ty.append(np.full(len(chain), t_c))
        n += len(chain)
        ms.append(n)
    for p in ions:
        xyz.append(np.asarray(p).reshape(1, 3))
        q.append([q_i])
        ty.append([t_i])
        n += 1
        ms.append(n)
    return dict(sd, xyz=np.concatenate(xyz), q=np.concatenate(q),
                type=np.concatenate(ty).astype(np.int32), mol_start=np.array(ms, dtype=np.int32))


def _nonbond(sd):
    o = O.Oracle(sd)
    e = o.energy_full()
    o.close()
    return sum(e[k] for k in NONBOND)


def _walk(rng, start, n, box, r0=1.0):
    out = [np.asarray(start, dtype=float)]
    while len(out) < n:
        v = rng.normal(size=3)
        out.append(out[-1] + r0 * v / np.linalg.norm(v))
    return np.array(out)


@pytest.mark.parametrize("name,charged,cur", [("small", True, 0), ("small", True, 1),
                                              ("small", True, 4), ("small", False, 3),
                                              ("gc_small", True, 5)])
def test_bead_energy_is_a_difference_of_full_energies(name, charged, cur):
    sd = systems.config(name)
    rng = np.random.default_rng(11 + cur)
    box = np.asarray(sd["box"])
    q_c, q_i = (-1.0, 1.0) if charged else (0.0, 0.0)
    chain = _walk(rng, rng.uniform(0, 1, 3) * box, cur + 1, box)
    ions = rng.uniform(0, 1, (cur + 1, 3)) * box
    grown, b1 = chain[:cur], chain[cur]
    gi, b2 = ions[:cur], ions[cur]
    o = O.Oracle(sd)
    got = o.cbmc_beads_energy(b1, 0, q_c, grown=grown, b2=b2 if charged else None, t2=1, q2=q_i,
                              grown_ions=gi if charged else None)
    o.close()
    sa = _with_chain(sd, grown, gi if charged else [], q_c, q_i)
    sb = _with_chain(sd, chain, ions if charged else [], q_c, q_i)
    ob = O.Oracle(sb)
    n0 = sd["mol_start"][-1]
    bonded = ob.pair("lj", n0 + cur, n0 + cur - 1) if cur > 0 else 0.0
    ob.close()
    want = _nonbond(sb) - _nonbond(sa) - bonded
    print(name, charged, cur, got, want, flush=True)
    assert abs(got - want) <= 1e-9 * max(1.0, abs(want))


def test_bead_energy_with_walls_is_a_difference_of_full_energies():
    sd = systems.config("slab_small", dipole=0)
    rng = np.random.default_rng(5)
    box = np.asarray(sd["box"])
    cur = 3
    chain = np.array([[4.0, 4.0, 3.0 + 0.9 * k] for k in range(cur + 1)])
    ions = np.c_[rng.uniform(0, 1, (cur + 1, 2)) * box[:2], rng.uniform(1.5, box[2] - 1.5, cur + 1)]
    o = O.Oracle(sd)
    got = o.cbmc_beads_energy(chain[cur], 0, -1.0, grown=chain[:cur], b2=ions[cur], t2=1, q2=1.0,
                              grown_ions=ions[:cur])
    o.close()
    sa = _with_chain(sd, chain[:cur], ions[:cur], -1.0, 1.0)
    sb = _with_chain(sd, chain, ions, -1.0, 1.0)
    ob = O.Oracle(sb)
    n0 = sd["mol_start"][-1]
    bonded = ob.pair("lj", n0 + cur, n0 + cur - 1)
    ob.close()
    want = _nonbond(sb) - _nonbond(sa) - bonded
    assert abs(got - want) <= 1e-9 * max(1.0, abs(want))


def test_excluded_molecules_are_left_out():
    sd = systems.config("gc_small")
    L = sd["gc"]["chain_len"]
    ms = sd["mol_start"]
    c = 2
    lo, hi = c * (1 + L), c * (1 + L) + L
    keep = np.r_[0:ms[lo], ms[hi + 1]:ms[-1]]
    kept_ms = np.r_[ms[:lo + 1], ms[hi + 2:] - (ms[hi + 1] - ms[lo])]
    sd2 = dict(sd, xyz=sd["xyz"][keep], q=sd["q"][keep], type=sd["type"][keep],
               mol_start=kept_ms.astype(np.int32))
    rng = np.random.default_rng(3)
    box = np.asarray(sd["box"])
    o1, o2 = O.Oracle(sd), O.Oracle(sd2)
    for _ in range(6):
        grown = _walk(rng, rng.uniform(0, 1, 3) * box, 3, box)
        gi = rng.uniform(0, 1, (3, 3)) * box
        b1 = grown[-1] + np.array([0.6, 0.8, 0.0])
        b2 = rng.uniform(0, 1, 3) * box
        e1 = o1.cbmc_beads_energy(b1, 0, -1.0, grown=grown, b2=b2, t2=1, q2=1.0, grown_ions=gi,
                                  exclude_mols=(lo, hi))
        e2 = o2.cbmc_beads_energy(b1, 0, -1.0, grown=grown, b2=b2, t2=1, q2=1.0, grown_ions=gi)
        assert abs(e1 - e2) <= 1e-10 * max(1.0, abs(e2))
    o1.close()
    o2.close()


def test_overlap_is_forbidden():
    sd = systems.config("small")
    o = O.Oracle(sd)
    p = np.asarray(sd["xyz"][5])
    assert o.cbmc_beads_energy(p + 1e-3, 0, -1.0, b2=np.array([1.0, 1.0, 1.0]), t2=1,
                               q2=1.0) == 1e8
    o.close()
