"""The wall-force slit pressure (ForceField::CalcPressureForceLJELSlit,
pressure.cc:404-470): the oracle's restatement pinned by what the mathematics
fixes, and the host implementation (plum_slit_forces) against the oracle.

Pins of the oracle:
- each pair force is minus the derivative of the oracle's own pair energy
  (PairForce = -dE_LJ/dr below the cutoff; PairForceZReal / PairForceZRepl =
  -dE/dz of the wall site for the real-space and reciprocal pair energies;
  BeadForceOnWall = -dE/dz near the z = 0 wall and +dE/dz near the z = Lz
  wall), by central differences;
- the assembled sums: translating the z = 0 wall's sites rigidly by dZ changes
  the total energy by -(forces on them) dZ, and likewise the other wall, so
  with D1, D2 the two derivatives of the oracle's total LJ (or real + k-space)
  energy, the sums satisfy f_ion + f_chain = D1/2 - D2/2 - f_wallwall, and
  f_wallwall = D1 of a system holding only the two walls (the -1/2, +1/2 and
  -1 of pressure.cc:425-429; a wrong sign, coefficient or partner class fails
  it).  The dipole term stays off: the reference's slit pressure has none.
The reference's own quirks are kept and checked as written: PairForce beyond
a full LJ cutoff uses (sigma / lj_cutoff)^6, and the grafted beads' FENE term
is R0 times the derivative of the FENE energy.
"""
import numpy as np
import pytest

from oracle import oracle as O
from plum_amd import host as H

H_FD = 1e-5
TYPES = [{"name": "P", "sigma": 1.0, "eps": 1.0, "wall_sigma": 1.0, "wall_eps": 1.0},
         {"name": "C", "sigma": 1.0, "eps": 1.0, "wall_sigma": 1.0, "wall_eps": 1.0},
         {"name": "Z", "sigma": 1.0, "eps": 1.0, "wall_sigma": 1.0, "wall_eps": 0.0},
         {"name": "L", "sigma": 1.0, "eps": 1.0, "wall_sigma": 1.0, "wall_eps": 1.0}]


def slit_system(seed=0, sites=4, chain=5, ions=6, use_pair=1, use_ewald=1, use_ext=0, lj_cutoff=-1.0,
                wall_cut=-1.0, system=True, zmin=1.0):
    """Phantom wall sites (type Z, charges +-1) on z = 0 and z = Lz first,
    then one chain (-1 per bead, its first bead grafted 'L' when use_ext) and
    ions (+1; two of them, like the chain's first bead, within LJ range of a
    site), random positions at least 0.95 apart."""
    rng = np.random.default_rng(seed)
    box = np.array([5.0, 5.0, 7.0])
    xyz, q, ty, ms = [], [], [], [0]

    def ok(p):
        return all(np.linalg.norm(np.asarray(o)[:2] - p[:2] - box[:2] * np.round((np.asarray(o)[:2] - p[:2]) / box[:2]))
                   ** 2 + (o[2] - p[2]) ** 2 > 0.95 ** 2 for o in xyz)

    for w, z in ((0, 0.0), (1, box[2])):
        for s in range(sites):
            xyz.append([(s % 2) * 2.5 + 0.3 * w + 0.1, (s // 2) * 2.5 + 0.2 * w + 0.3, z])
            q.append(1.0 if (s + w) % 2 == 0 else -1.0)
            ty.append(2)
            ms.append(len(xyz))
    if system:
        # the chain's first bead and two ions within LJ range of a site
        p = np.array([2.5, 2.7, 1.05])
        for b in range(chain):
            for _ in range(1000):
                v = rng.normal(size=3)
                cand = p + v / np.linalg.norm(v) if b > 0 else p
                if zmin < cand[2] < box[2] - zmin and ok(cand):
                    break
            p = cand
            xyz.append(list(p))
            q.append(-1.0)
            ty.append(3 if (b == 0 and use_ext) else 0)
        ms.append(len(xyz))
        for near in ([0.15, 0.35, 1.03], [0.45, 0.55, box[2] - 1.04]):
            xyz.append(near)
            q.append(1.0)
            ty.append(1)
            ms.append(len(xyz))
        for _ in range(ions):
            while True:
                cand = np.array([rng.uniform(0, box[0]), rng.uniform(0, box[1]), rng.uniform(zmin, box[2] - zmin)])
                if ok(cand):
                    break
            xyz.append(list(cand))
            q.append(1.0)
            ty.append(1)
            ms.append(len(xyz))
    return {"npbc": 2, "beta": 1.0, "use_pair": use_pair, "use_ewald": use_ewald, "use_bond": 0,
            "use_ext": use_ext, "lj_cutoff": lj_cutoff, "lB": 1.0, "dielectric": 80.0, "dipole": 0,
            "k_bond": 400.0, "r0": 1.0, "wall_cut": wall_cut, "wall_sigma_w": 1.0, "wall_eps_w": 1.0,
            "types": TYPES, "box": list(box), "alpha": 0.6, "xyz": np.array(xyz), "q": np.array(q),
            "type": np.array(ty, dtype=np.int32), "mol_start": np.array(ms, dtype=np.int32)}, 2 * sites


def _energy(sd, terms):
    e = O.Oracle(sd).energy_full()
    return sum(e[t] for t in terms)


def _shift(sd, rows, dz):
    s2 = dict(sd)
    x = np.array(sd["xyz"], dtype=float).copy()
    x[rows, 2] += dz
    s2["xyz"] = x
    return s2


def _d_dz(sd, rows, terms):
    return (_energy(_shift(sd, rows, H_FD), terms) - _energy(_shift(sd, rows, -H_FD), terms)) / (2 * H_FD)


@pytest.mark.parametrize("kind,terms,f_idx", [("lj", ("lj",), 0), ("ewald", ("real", "kspace"), 3)])
def test_slit_force_sums_are_minus_the_wall_translation_derivatives(kind, terms, f_idx):
    use_pair, use_ewald = (1, 0) if kind == "lj" else (0, 1)
    sd, ph = slit_system(seed=3, use_pair=use_pair, use_ewald=use_ewald, zmin=0.3, ions=10)
    f = O.Oracle(sd).slit_forces(ph)
    w1, w2 = list(range(ph // 2)), list(range(ph // 2, ph))
    d1, d2 = _d_dz(sd, w1, terms), _d_dz(sd, w2, terms)
    walls, _ = slit_system(seed=3, use_pair=use_pair, use_ewald=use_ewald, system=False, zmin=0.3)
    fw = O.Oracle(walls).slit_forces(ph)
    dw = _d_dz(walls, w1, terms)
    scale = max(abs(d1), abs(d2), abs(dw), 1.0)
    print(kind, f, d1, d2, fw[f_idx + 2], dw)
    assert abs(f[f_idx + 2] - fw[f_idx + 2]) <= 1e-12 * scale
    assert abs(fw[f_idx + 2] - dw) <= 1e-6 * scale
    assert abs(f[f_idx] + f[f_idx + 1] - (0.5 * d1 - 0.5 * d2 - f[f_idx + 2])) <= 1e-6 * scale
    assert abs(f[f_idx]) > 1e-3 and abs(f[f_idx + 1]) > 1e-3   # both partner classes take part
    other = 3 - f_idx
    assert np.all(f[other:other + 3] == 0)


def test_pair_forces_are_minus_the_pair_energy_derivatives():
    sd, ph = slit_system(seed=5, lj_cutoff=2.5)
    o = O.Oracle(sd)
    n = len(sd["q"])
    checked = 0
    for a in range(ph):
        for b in range(ph, n):
            for kind in ("real", "repl"):
                x = np.array(sd["xyz"], dtype=float)
                e = []
                for dz in (H_FD, -H_FD):
                    y = x.copy()
                    y[a, 2] += dz
                    oo = O.Oracle(dict(sd, xyz=y), init_energy=False)
                    e.append(oo.pair(kind, a, b))
                fd = -(e[0] - e[1]) / (2 * H_FD)
                got = o.force(kind, a, b)
                assert abs(got - fd) <= 1e-6 * max(1.0, abs(fd)), (a, b, kind, got, fd)
                checked += 1
    assert checked > 0
    # LJ: -dE/dr below the cutoff along the pair axis; beyond it the
    # reference's expression with (sigma / lj_cutoff)^6
    for r in (0.95, 1.1, 1.7, 2.4):
        two = dict(sd, xyz=np.array([[1.0, 1.0, 2.0], [1.0, 1.0, 2.0 + r]] + [[0, 0, 5.0]] * (n - 2)))
        e = []
        for dr in (H_FD, -H_FD):
            y = np.array(two["xyz"], dtype=float)
            y[1, 2] += dr
            e.append(O.Oracle(dict(two, xyz=y), init_energy=False).pair("lj", 0, 1))
        fd = -(e[0] - e[1]) / (2 * H_FD)
        got = O.Oracle(two, init_energy=False).force("lj", 0, 1)
        assert abs(got - fd) <= 1e-6 * max(1.0, abs(fd)), (r, got, fd)
    r = 3.0
    two = dict(sd, xyz=np.array([[1.0, 1.0, 2.0], [1.0, 1.0, 2.0 + r]] + [[0, 0, 5.0]] * (n - 2)))
    r6 = (1.0 / 2.5) ** 6
    assert O.Oracle(two, init_energy=False).force("lj", 0, 1) == pytest.approx(4 * (r6 * r6 * 12 / r - r6 * 6 / r),
                                                                              rel=1e-12)


@pytest.mark.parametrize("z,sign", [(0.9, -1.0), (1.05, -1.0), (6.1, 1.0), (5.95, 1.0)])
def test_wall_force_is_the_wall_energy_derivative(z, sign):
    sd, ph = slit_system(seed=1, use_pair=0, use_ewald=0, use_ext=1)
    n = len(sd["q"])
    k = ph + 1   # the chain's second bead: a plain wall type
    x = np.array(sd["xyz"], dtype=float)
    x[k, 2] = z
    e = []
    for dz in (H_FD, -H_FD):
        y = x.copy()
        y[k, 2] += dz
        e.append(O.Oracle(dict(sd, xyz=y), init_energy=False).wall(k))
    fd = (e[0] - e[1]) / (2 * H_FD)
    got = O.Oracle(dict(sd, xyz=x), init_energy=False).force("wall", k)
    assert abs(got - sign * fd) <= 1e-6 * max(1.0, abs(fd)), (z, got, fd)
    assert n > k


def test_grafted_wall_force_keeps_the_reference_fene_term():
    sd, ph = slit_system(seed=1, use_pair=0, use_ewald=0, use_ext=1)
    k = ph   # the grafted 'L' bead
    x = np.array(sd["xyz"], dtype=float)
    z = 0.8
    x[k, 2] = z
    got = O.Oracle(dict(sd, xyz=x), init_energy=False).force("wall", k)
    r3 = (1.0 / z) ** 3
    R0 = 3 * 1.25992104989
    want = 2.59807621135 * (r3 * r3 * 6 / z - r3 * 3 / z) - R0 * z / (1 - (z / R0) ** 2)
    assert got == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("seed,lj_cutoff,use_ext,wall_cut", [(0, -1.0, 0, -1.0), (1, 2.5, 0, -1.0),
                                                              (2, -1.0, 1, -1.0), (4, 2.5, 1, 1.5)])
def test_host_slit_forces_match_the_oracle(seed, lj_cutoff, use_ext, wall_cut):
    sd, ph = slit_system(seed=seed, lj_cutoff=lj_cutoff, use_ext=use_ext, wall_cut=wall_cut, zmin=0.4)
    o = O.Oracle(sd)
    ref = o.slit_forces(ph)
    got = H.slit_forces(sd, o.ewald_info(), ph)
    print(ref, got)
    scale = max(1.0, np.abs(ref).max())
    assert np.all(np.abs(got - ref) <= 1e-10 * scale), (got, ref)
