"""The volume-perturbation pressure estimator of the oracle
(ForceField::CalcPressureVolScalingHSELSlit, pressure.cc:187-387) pinned
without the reference: the stretch energy dU of a bulk system equals the
difference of two from-scratch energies (the system with every z scaled by
1 + dz/Lz in a box longer by dz, minus the system as it is), whose code path
shares only the pair functions; rigidly shifted molecules leave their
intramolecular energy unchanged; and with every interaction off the
estimators reduce to the ideal gas."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from plum_amd import systems

KDZ = 1e-5


def _stretched(sd):
    """The same system with z scaled by 1 + dz/Lz in a box of Lz + dz."""
    L = sd["box"]
    xyz = np.array(sd["xyz"], dtype=float).copy()
    xyz[:, 2] *= 1.0 + KDZ / L[2]
    return dict(sd, xyz=xyz, box=[L[0], L[1], L[2] + KDZ])


@pytest.mark.parametrize("name", ["tiny", "small"])
def test_stretch_energy_is_a_difference_of_full_energies(name):
    sd = systems.config(name)
    o = O.Oracle(sd)
    s = o.vp_sample()
    e0 = o.energy_full()
    o2 = O.Oracle(_stretched(sd))
    assert list(o2.ewald_info().repl_cell) == list(o.ewald_info().repl_cell)
    e1 = o2.energy_full()
    parts = {"el": e1["real"] + e1["kspace"] - e0["real"] - e0["kspace"],
             "lj": e1["lj"] - e0["lj"], "bond": e1["bond"] - e0["bond"]}
    scale = sum(abs(v) for v in e0.values())
    for k, v in parts.items():
        assert abs(s[k] - v) <= 1e-10 * scale, (k, s[k], v)
    assert abs(s["dU"] - sum(parts.values())) <= 1e-10 * scale
    assert s["n_mol"] == len(sd["mol_start"]) - 1
    assert s["vol"] == pytest.approx(np.prod(sd["box"]), rel=1e-15)
    # the stretch changes the energy by O(dz): dU / dz is a finite force
    assert 0 < abs(s["dU"]) < 1e3 * KDZ * scale
    o.close()
    o2.close()


def test_rigid_bonds_shift_molecules_whole():
    """Without the spring potential every molecule moves by kDz com_z/Lz as a
    whole: its internal LJ and real-space pairs do not change, so a single
    isolated molecule (no partners within reach) has dU from its periodic
    images and k space only."""
    sd = systems.config("tiny")
    sd = dict(sd, use_bond=0)
    o = O.Oracle(sd)
    s = o.vp_sample()
    # only one molecule's beads, far from every image: LJ part is zero
    ms = sd["mol_start"]
    one = dict(sd, xyz=np.array(sd["xyz"])[ms[0]:ms[1]], q=np.array(sd["q"])[ms[0]:ms[1]],
               type=np.array(sd["type"])[ms[0]:ms[1]], mol_start=np.array([0, ms[1] - ms[0]],
                                                                          dtype=np.int32))
    o1 = O.Oracle(dict(one, box=[40.0, 40.0, 40.0], alpha=(5.0 / 40) ** 2))
    s1 = o1.vp_sample()
    assert s1["lj"] == 0.0 and s1["bond"] == 0.0
    o.close()
    o1.close()


def test_ideal_gas_limit():
    """No interactions: dU = 0, so the de Miguel estimator is
    ln((1 + dz/Lz)^N) / (A dz), the ideal N/V to O(dz), and the Yethiraj
    estimator is N/V exactly."""
    sd = systems.config("small")
    sd = dict(sd, use_pair=0, use_ewald=0, use_bond=0)
    o = O.Oracle(sd)
    s = o.vp_sample()
    assert s["dU"] == 0.0
    L = sd["box"]
    n = s["n_mol"]
    p_dm = math.log((1 + KDZ / L[2]) ** n) / s["a_dz"]
    assert p_dm == pytest.approx(n / s["vol"], rel=1e-4)
    o.close()
