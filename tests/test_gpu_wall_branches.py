"""Every branch of the 6-3 wall on the GPU against the oracle
(potential_truncated_lj_wall.cc:46-134): plain beads with the repulsive-only
wall (cut < 0) and with the cut full well (cut >= 0), left-grafted "L" and
right-grafted "R" beads with their FENE tethers, near and far from each wall;
full energies and per-move dE."""
import numpy as np
import pytest

from oracle import oracle as O
from plum_amd import engine as E

pytestmark = pytest.mark.gpu


def _sys(wall_cut, Lz=6.0):
    types = [{"name": "P", "sigma": 1.0, "eps": 1.0, "wall_sigma": 1.0, "wall_eps": 2.0},
             {"name": "L", "sigma": 1.0, "eps": 1.0, "wall_sigma": 1.3, "wall_eps": 1.5},
             {"name": "R", "sigma": 1.0, "eps": 1.0, "wall_sigma": 0.9, "wall_eps": 1.0}]
    zs = [0.6, 1.0, 1.4, 3.0, 4.7, 5.2, 0.9, 1.6, 2.5, 5.1, 4.3, 3.6]
    ty = [0, 0, 0, 0, 0, 0, 1, 1, 1, 2, 2, 2]
    n = len(zs)
    xyz = np.array([[1.0 + 2.5 * (i % 4), 1.0 + 2.5 * (i // 4), z] for i, z in enumerate(zs)])
    return dict(box=[12.0, 12.0, Lz], npbc=2, beta=1.0, use_pair=0, use_ewald=0, use_bond=0,
                use_ext=1, lj_cutoff=2.5, types=types, xyz=xyz, q=np.zeros(n),
                type=np.array(ty, dtype=np.int32), mol_start=np.arange(n + 1, dtype=np.int32),
                wall_cut=wall_cut, move_prob=[1.0, 0, 0, 0, 0], move_size=0.3)


@pytest.mark.parametrize("wall_cut", [-1.0, 1.5, 2.5])
def test_wall_branches_match_oracle(wall_cut):
    sd = _sys(wall_cut)
    eng = E.Engine(sd)
    orc = O.Oracle(sd)
    got, ref = eng.energy_full()["wall"], orc.energy_full()["wall"]
    assert abs(got - ref) <= 1e-10 * max(abs(ref), 1e-2), (got, ref)
    rng = np.random.default_rng(5)
    for i in range(len(sd["q"])):
        tr = sd["xyz"][i] + np.array([0.0, 0.0, rng.uniform(-0.35, 0.35)])
        t_g, d_g = eng.delta(i, [i], tr[None, :])
        orc.set_trial(i, [i], tr[None, :])
        t_r, d_r = orc.delta(i)
        assert (t_g >= 1e8) == (t_r >= 1e8), i
        if t_r < 1e8:
            assert abs(d_g["wall"] - d_r["wall"]) <= 1e-10 * max(abs(d_r["wall"]), 1e-2), (i, d_g, d_r)
        eng.finalize(True)
        orc.finalize(i, True)
    assert np.array_equal(eng.positions(), orc.positions())
