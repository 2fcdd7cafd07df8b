"""The volume-perturbation pressure sample on the GPU (plum_vp_sample)
against the oracle's restatement of pressure.cc:187-387, term by term, on
bulk and slab systems, with and without bonds and surface sites.

Tolerance: dU is a sum of pair-energy changes (new minus old per pair, and
in k space c'_k |S'|^2 - c_k |S|^2) of energies of total magnitude
E = sum |E_term|; the reference rounds each coefficient c_k, c'_k and each
pair sum on its own, which leaves ~1e-16 E of noise in its dU.  The GPU and
the oracle agree to 1e-14 E (a hundred times that floor); with dz = 1e-5 this
is ~1e-9 of dU, and 1e-9 of the pressure dU / (A dz)."""
import numpy as np
import pytest

from oracle import oracle as O
from plum_amd import engine as E
from plum_amd import systems

pytestmark = pytest.mark.gpu
PARTS = ("dU", "el", "lj", "wall", "bond", "dipole")


def _with_sites(sd, n_sites=8):
    """sd with n_sites charged single-bead surface sites (no LJ, no wall)
    put first, half near z = 0 and half near z = Lz: the phantom sites the
    pair rule of pressure.cc:271-272 singles out."""
    L = sd["box"]
    rng = np.random.default_rng(11)
    h = n_sites // 2
    z = np.r_[np.full(h, 0.3), np.full(n_sites - h, L[2] - 0.3)]
    sites = np.c_[rng.uniform(0, L[0], n_sites), rng.uniform(0, L[1], n_sites), z]
    qs = np.r_[np.ones(h), -np.ones(n_sites - h)]
    types = list(sd["types"]) + [{"name": "Z", "sigma": 1.0, "eps": 0.0, "wall_sigma": 1.0,
                                  "wall_eps": 0.0}]
    tz = len(types) - 1
    ms = np.r_[np.arange(n_sites), np.asarray(sd["mol_start"]) + n_sites].astype(np.int32)
    return dict(sd, xyz=np.r_[sites, sd["xyz"]], q=np.r_[qs, sd["q"]],
                type=np.r_[np.full(n_sites, tz), sd["type"]].astype(np.int32), mol_start=ms,
                types=types, phantom=n_sites), n_sites


def _sys(name):
    if name == "small_rigid":
        return dict(systems.config("small"), use_bond=0), 0
    if name == "slab_sites":
        return _with_sites(systems.config("slab_small"))
    sd = systems.config(name)
    return sd, sd.get("phantom", 0)


@pytest.mark.parametrize("name", ["tiny", "small", "small_rigid", "slab_small", "slab_sites", "c0"])
def test_vp_sample_matches_oracle(name):
    sd, phantom = _sys(name)
    eng = E.Engine(sd)
    orc = O.Oracle(sd)
    got, ref = eng.vp_sample(phantom), orc.vp_sample(phantom)
    scale = sum(abs(v) for v in orc.energy_full().values())
    print(name, "E %.3g" % scale, " ".join("%s: %.3e rel %.1e" % (
        k, ref[k], abs(got[k] - ref[k]) / max(abs(ref[k]), 1e-300)) for k in PARTS), flush=True)
    for k in PARTS:
        assert abs(got[k] - ref[k]) <= 1e-14 * scale, (k, got[k], ref[k])
    assert got["n_mol"] == ref["n_mol"]
    assert got["vol"] == pytest.approx(ref["vol"], rel=1e-15)
    assert got["area_dz"] == pytest.approx(ref["a_dz"], rel=1e-15)
    # after device moves the sample follows the new configuration
    eng.mc_run(300, sd["move_prob"], move_size=sd["move_size"], seed=3,
               phantom=sd.get("phantom", 0), grafted=sd.get("grafted", 0), bond_len=1.0)
    orc2 = O.Oracle(dict(sd, xyz=eng.positions()))
    got, ref = eng.vp_sample(phantom), orc2.vp_sample(phantom)
    for k in PARTS:
        assert abs(got[k] - ref[k]) <= 1e-14 * scale, (k, got[k], ref[k])


def test_host_pressure_estimators_follow_the_reference_accumulation(tmp_path):
    """The host driver's pressure (ForceField::CalcPressureVolScalingHSELSlit
    sampled every 40 steps along the reference's own move sequence) equals
    the reference formulas applied to the oracle's samples on the oracle's
    replay of the same sequence: Yethiraj beta <N/V> - <dU>/(A dz) and de
    Miguel beta/(A dz) ln <(1 + dz/Lz)^N e^{-beta dU}>."""
    import math
    from plum_amd import host as H
    sd = systems.config("small")
    path = systems.write_reference_inputs(sd, str(tmp_path))
    sim = H.HostSimulation(path, 5)
    orc = O.Oracle(sd)
    rng = O.MT19937(5)
    ideal = du = dm = 0.0
    for k in range(6):
        sim.run(40)
        for _ in range(40):
            O.mc_step(orc, rng, sd["move_prob"], sd["move_size"], 0, 0, sd["r0"], True)
        v, py, pd, n = sim.pressure_sample()
        r = orc.vp_sample(0)
        assert abs(v["dU"] - r["dU"]) <= 1e-14 * 300.0, (v["dU"], r["dU"])
        ideal += r["n_mol"] / r["vol"]
        du += r["dU"]
        dm += (1 + 1e-5 / sd["box"][2]) ** r["n_mol"] * math.exp(-r["dU"])
        py_ref = (ideal - du / r["a_dz"]) / (k + 1)
        pd_ref = 1.0 / r["a_dz"] * math.log(dm / (k + 1))
        assert n == k + 1
        assert py == pytest.approx(py_ref, rel=1e-9, abs=1e-12)
        assert pd == pytest.approx(pd_ref, rel=1e-7, abs=1e-9)
    print("small: P_Yethiraj %.6f P_deMiguel %.6f (ideal %.6f)" % (py, pd, ideal / 6), flush=True)
    orc.close()
    sim.close()
