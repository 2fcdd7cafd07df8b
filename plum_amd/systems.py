"""Seeded synthetic systems in the reference's data model, and writers for the
reference input format (run.in + crd + top).

A system definition ("sysdef") is a dict:
  box[3], npbc, beta, use_pair/use_ewald/use_bond/use_ext,
  types: list of {name, sigma, eps, wall_sigma, wall_eps}
  xyz (N,3), q (N,), type (N,), mol_start (M+1,), phantom, grafted,
  lj_cutoff, lB, dielectric, alpha (kappa^2, sigma^-2), dipole,
  k_bond, r0, wall_cut, move_prob[5], move_size.

Chains are random walks with step r0 along Marsaglia directions; a step (or
an ion) is rejected when it lands within `min_sep` of an existing bead, which
keeps LJ finite.  Molecule order follows the reference's convention
(simulation.cc:193-199, cbmc.cc:343-346): phantom surface sites first, then
chains, then ions.
"""
import math
import os

import numpy as np

CONFIG_SEEDS = {"c0": 1000, "c1": 1001, "c2": 1002, "c3": 1003, "c4": 1004}
# chemical potential (kT) of the exchanged chain + counterions of the
# grand-canonical workload, chosen where the 50 starting chains stay in
# balance.  Runs from the start (one exchange every 10 steps) end at 3
# chains for mu = -185 and 16 for -175 after 8000 steps; over 20000 steps
# -168 holds 51 with one seed and drains to 31 with another, while -165
# settles at 54 and 44 (second-half means equal to the first within the
# fluctuations, ~20 insertions and ~20 deletions accepted each way).  Widom
# test insertions do not locate it for this exchange (a 32-bead chain and 32
# counterions grown at random): their weights spread over ~85 kT between
# batches of 50, and the mean is set by the rare largest ones.
GC_CHEM_POT = -165.0


class _Grid:
    """Uniform hash grid for the overlap rejection of the generator."""

    def __init__(self, box, h, periodic):
        self.box = np.asarray(box, dtype=float)
        self.h = h
        self.periodic = periodic
        self.n = np.maximum(1, (self.box // h).astype(int))
        self.cells = {}

    def _key(self, p):
        c = np.floor(self._wrap(p) / (self.box / self.n)).astype(int)
        return tuple(np.clip(c, 0, self.n - 1))

    def _wrap(self, p):
        p = np.array(p, dtype=float)
        for d in range(3):
            if self.periodic[d]:
                p[d] -= self.box[d] * math.floor(p[d] / self.box[d])
        return p

    def add(self, p):
        self.cells.setdefault(self._key(p), []).append(np.array(p, dtype=float))

    def too_close(self, p, r):
        k = self._key(p)
        r2 = r * r
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    c = [k[0] + dx, k[1] + dy, k[2] + dz]
                    ok = True
                    for d in range(3):
                        if self.periodic[d]:
                            c[d] %= self.n[d]
                        elif c[d] < 0 or c[d] >= self.n[d]:
                            ok = False
                    if not ok:
                        continue
                    for q in self.cells.get(tuple(c), ()):
                        d = p - q
                        for e in range(3):
                            if self.periodic[e]:
                                d[e] -= self.box[e] * round(d[e] / self.box[e])
                        if d @ d < r2:
                            return True
        return False


def _sphere(rng):
    while True:
        r1, r2 = 1 - 2 * rng.random(), 1 - 2 * rng.random()
        s = r1 * r1 + r2 * r2
        if s <= 1:
            h = 2 * math.sqrt(1 - s)
            return np.array([r1 * h, r2 * h, 1 - 2 * s])


def _std_types(extra=()):
    t = [{"name": "P", "sigma": 1.0, "eps": 1.0, "wall_sigma": 1.0, "wall_eps": 1.0},
         {"name": "C", "sigma": 1.0, "eps": 1.0, "wall_sigma": 1.0, "wall_eps": 1.0}]
    t.extend(extra)
    return t


def build_bulk(n_chains, chain_len, n_ions, box_l, seed, bead_q=-1.0, ion_q=1.0,
               n_salt_pairs=0, min_sep=0.8, r0=1.0, z_range=None, npbc=3,
               graft=None, phantom_sites=None, box=None):
    """Random-walk chains + uniform ions.  Returns (xyz, q, type, mol_start)."""
    rng = np.random.default_rng(seed)
    if box is None:
        box = [box_l, box_l, box_l]
    periodic = [True, True, npbc == 3]
    grid = _Grid(box, max(min_sep, 1.0), periodic)
    xyz, q, ty, ms = [], [], [], [0]
    zlo, zhi = (0.0, box[2]) if z_range is None else z_range

    def uniform_point():
        return np.array([rng.random() * box[0], rng.random() * box[1],
                         zlo + rng.random() * (zhi - zlo)])

    if phantom_sites is not None:
        for (p, qq, t) in phantom_sites:
            xyz.append(np.asarray(p, dtype=float)); q.append(qq); ty.append(t)
            grid.add(np.asarray(p, dtype=float))
            ms.append(len(xyz))
    for c in range(n_chains):
        while True:
            if graft is not None:
                start = np.array([rng.random() * box[0], rng.random() * box[1], graft])
            else:
                start = uniform_point()
            if not grid.too_close(start, min_sep):
                break
        chain = [start]
        grid.add(start)
        tries = 0
        while len(chain) < chain_len:
            step = _sphere(rng) * r0
            p = chain[-1] + step
            if z_range is not None and not (zlo < p[2] < zhi):
                tries += 1
                if tries > 10000:
                    raise RuntimeError("chain growth failed")
                continue
            if grid.too_close(p, min_sep):
                tries += 1
                if tries > 10000:
                    raise RuntimeError("chain growth failed")
                continue
            chain.append(p)
            grid.add(p)
        for k, p in enumerate(chain):
            xyz.append(p)
            q.append(bead_q)
            ty.append(2 if (graft is not None and k == 0) else 0)
        ms.append(len(xyz))
    ions = [(ion_q, 1)] * n_ions + [(1.0, 1), (-1.0, 1)] * n_salt_pairs
    for (qq, t) in ions:
        while True:
            p = uniform_point()
            if not grid.too_close(p, min_sep):
                break
        grid.add(p)
        xyz.append(p); q.append(qq); ty.append(t)
        ms.append(len(xyz))
    return (np.array(xyz, dtype=np.float64).reshape(-1, 3), np.array(q, dtype=np.float64),
            np.array(ty, dtype=np.int32), np.array(ms, dtype=np.int32))


def gc_layout(xyz, q, ty, ms, n_chains, chain_len):
    """Reorder a bulk system of n_chains chains followed by n_chains *
    chain_len single-bead counterions into the grand-canonical layout of the
    reference (cbmc.cc:367-373: each exchanged chain is followed by its
    chain_len counterions, so chain c is molecule c (1 + chain_len))."""
    order = []
    n_ch_mols = n_chains
    for c in range(n_chains):
        order.append(c)
        order.extend(n_ch_mols + c * chain_len + k for k in range(chain_len))
    nx, nq, nt, nms = [], [], [], [0]
    for m in order:
        a, b = ms[m], ms[m + 1]
        nx.append(xyz[a:b]); nq.append(q[a:b]); nt.append(ty[a:b])
        nms.append(nms[-1] + b - a)
    return (np.concatenate(nx), np.concatenate(nq), np.concatenate(nt).astype(np.int32),
            np.array(nms, dtype=np.int32))


def config(name, seed=None, **over):
    """Named workloads.  c0/c1/c3 follow the stated configurations; 'tiny' and
    'small' are quick test systems."""
    name = name.lower()
    seed = CONFIG_SEEDS.get(name, 7) if seed is None else seed
    base = dict(npbc=3, beta=1.0, use_pair=1, use_ewald=1, use_bond=1, use_ext=0,
                lj_cutoff=2.5, lB=1.0, dielectric=80.0, dipole=0, k_bond=400.0, r0=1.0,
                wall_cut=-1.0, wall_sigma_w=1.0, wall_eps_w=1.0, phantom=0, grafted=0,
                move_prob=[0.0, 0.1, 0.4, 0.3, 0.2], move_size=1.0, types=_std_types())
    if name == "c0":
        L = (1000 / 0.1) ** (1 / 3)
        xyz, q, ty, ms = build_bulk(20, 25, 500, L, seed)
        base.update(box=[L, L, L], alpha=(5.0 / L) ** 2)
    elif name == "c1":
        L = (40000 / 0.05) ** (1 / 3)
        xyz, q, ty, ms = build_bulk(200, 100, 20000, L, seed)
        base.update(box=[L, L, L], alpha=0.08)
    elif name == "c3":
        n_ch, clen, lz = 64, 50, 40.0
        lxy = math.sqrt(n_ch / 0.1)
        rng = np.random.default_rng(seed + 17)
        sites = []
        qs = np.array([1.0] * 200 + [-1.0] * 200)
        rng.shuffle(qs)
        for k in range(400):
            sites.append(([rng.random() * lxy, rng.random() * lxy, lz], qs[k], 3))
        types = _std_types([{"name": "L", "sigma": 1.0, "eps": 1.0, "wall_sigma": 1.0,
                             "wall_eps": 1.0},
                            {"name": "Z", "sigma": 1.0, "eps": 0.0, "wall_sigma": 1.0,
                             "wall_eps": 0.0}])
        # grafted chains + counterions from the native generator (chains that
        # walk into a dead end are regrown), surface sites first
        from . import host
        cx, cq, cty, cms = host.synth_system(n_ch, clen, n_ch * clen, 0, [lxy, lxy, lz], seed,
                                             npbc=2, graft_z=1.0, z_range=(0.8, lz - 0.8))
        sx = np.array([p for (p, _, _) in sites])
        xyz = np.concatenate([sx, cx])
        q = np.concatenate([[qq for (_, qq, _) in sites], cq])
        ty = np.concatenate([[t for (_, _, t) in sites], cty]).astype(np.int32)
        ms = np.concatenate([np.arange(len(sites)), cms + len(sites)]).astype(np.int32)
        base.update(box=[lxy, lxy, lz], npbc=2, alpha=(5.0 / lxy) ** 2, dipole=1,
                    use_ext=1, types=types, phantom=400, grafted=n_ch,
                    move_prob=[0.0, 0.2, 0.5, 0.3, 0.0])
    elif name == "c4":
        L, n_ch, clen = 60.0, 50, 32
        xyz, q, ty, ms = gc_layout(*build_bulk(n_ch, clen, n_ch * clen, L, seed), n_ch, clen)
        base.update(box=[L, L, L], alpha=(5.0 / L) ** 2, gc=dict(
            chain_len=clen, bead_charge=-1.0, trials=16, freq=10, chem_pot=GC_CHEM_POT,
            debroglie=1.0))
    elif name == "gc_small":
        L, n_ch, clen = 12.0, 4, 6
        xyz, q, ty, ms = gc_layout(*build_bulk(n_ch, clen, n_ch * clen, L, seed), n_ch, clen)
        base.update(box=[L, L, L], alpha=(5.0 / L) ** 2, gc=dict(
            chain_len=clen, bead_charge=-1.0, trials=8, freq=4, chem_pot=0.0, debroglie=1.0))
    elif name == "tiny":
        L = 8.0
        xyz, q, ty, ms = build_bulk(2, 5, 10, L, seed)
        base.update(box=[L, L, L], alpha=(5.0 / L) ** 2)
    elif name == "small":
        L = 14.0
        xyz, q, ty, ms = build_bulk(6, 12, 72, L, seed)
        base.update(box=[L, L, L], alpha=(5.0 / L) ** 2)
    elif name == "slab_small":
        lxy, lz = 10.0, 12.0
        types = _std_types([{"name": "L", "sigma": 1.0, "eps": 1.0, "wall_sigma": 1.0,
                             "wall_eps": 1.0}])
        xyz, q, ty, ms = build_bulk(4, 8, 32, None, seed, box=[lxy, lxy, lz],
                                    z_range=(0.8, lz - 0.8), npbc=2, graft=1.0)
        base.update(box=[lxy, lxy, lz], npbc=2, alpha=(5.0 / lxy) ** 2, dipole=1,
                    use_ext=1, types=types, grafted=4, move_prob=[0.0, 0.2, 0.5, 0.3, 0.0])
    else:
        raise ValueError("unknown config " + name)
    base.update(xyz=xyz, q=q, type=ty, mol_start=ms, name=name)
    base.update(over)
    return base


_M64 = (1 << 64) - 1


def splitmix64(x):
    """SplitMix64 output for state x (Steele, Lea, Flood 2014)."""
    z = (x + 0x9E3779B97F4A7C15) & _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def alpha_for_cutoff(rc, lB=1.0):
    """Ewald alpha (kappa^2) for which the reference's cutoff rule
    (potential_ewald_coul.cc:56-61: smallest integer r with
    lB/2 erfc(sqrt(alpha) r)/r <= 1e-7) lands exactly on the integer rc."""
    target = 2e-7 * rc / lB
    lo, hi = 0.0, 20.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if math.erfc(mid) > target:
            lo = mid
        else:
            hi = mid
    return (hi / rc) ** 2 * (1 + 1e-9)


def replica_spec(r, seed=1002, n_max=None):
    """Draw replica r of the batched-replica workload: chain length U{10..200},
    chain count U{5..100}, added 1:1 salt U{0..2000} ion pairs, bead charge -1
    or 0 with p = 0.5 (charged chains carry one +1 counterion per bead),
    density U[0.02, 0.2].  The real-space cutoff is fixed by the builder at
    min(12, floor(L/2)) (at least 3) and alpha follows from it."""
    rng = np.random.default_rng(splitmix64((seed << 20) + r))
    clen = int(rng.integers(10, 201))
    nch = int(rng.integers(5, 101))
    salt = int(rng.integers(0, 2001))
    charged = bool(rng.random() < 0.5)
    rho = float(rng.uniform(0.02, 0.2))
    if n_max is not None:
        # reduced copies for tests: shrink chains and salt in proportion
        while nch * clen * (2 if charged else 1) + 2 * salt > n_max:
            nch = max(1, nch // 2)
            salt //= 2
            if nch == 1 and salt == 0 and clen > 10:
                clen = max(10, clen // 2)
            if nch == 1 and salt == 0 and clen == 10:
                break
    n = nch * clen * (2 if charged else 1) + 2 * salt
    L = (n / rho) ** (1.0 / 3.0)
    rc = int(max(3, min(12, math.floor(L / 2))))
    return dict(chain_len=clen, n_chains=nch, salt_pairs=salt, charged=charged, rho=rho,
                n=n, box=L, rc=rc, seed=splitmix64((seed << 20) + r + 0x5EED))


def replica_system(spec):
    """Sysdef of one replica, built by the native generator."""
    from . import host
    L = spec["box"]
    nci = spec["n_chains"] * spec["chain_len"] if spec["charged"] else 0
    xyz, q, ty, ms = host.synth_system(spec["n_chains"], spec["chain_len"], nci,
                                       spec["salt_pairs"], [L, L, L], spec["seed"] & _M64,
                                       bead_q=-1.0 if spec["charged"] else 0.0)
    return dict(npbc=3, beta=1.0, use_pair=1, use_ewald=1, use_bond=1, use_ext=0,
                lj_cutoff=2.5, lB=1.0, dielectric=80.0, dipole=0, k_bond=400.0, r0=1.0,
                wall_cut=-1.0, wall_sigma_w=1.0, wall_eps_w=1.0, phantom=0, grafted=0,
                move_prob=[0.0, 0.1, 0.4, 0.3, 0.2], move_size=1.0, types=_std_types(),
                box=[L, L, L], alpha=alpha_for_cutoff(spec["rc"]), xyz=xyz, q=q, type=ty,
                mol_start=ms, name="replica")


def replica_set(n=256, seed=1002, n_max=None):
    """The batched-replica workload: n independent bulk systems."""
    return [replica_system(replica_spec(r, seed, n_max)) for r in range(n)]


def write_reference_inputs(sysdef, outdir, steps=1000, prefix="output"):
    """Write run.in / input_crd.dat / input_top.dat in the reference format
    (Simulation ctor simulation.cc:30-50, ReadTop :610-646, ReadCrd :582-608,
    ForceField::Initialize force_field.cc:65-195 and the ReadParameters of
    each potential)."""
    os.makedirs(outdir, exist_ok=True)
    xyz, q, ty, ms = sysdef["xyz"], sysdef["q"], sysdef["type"], sysdef["mol_start"]
    names = [t["name"] for t in sysdef["types"]]
    n, m = len(q), len(ms) - 1
    with open(os.path.join(outdir, "input_crd.dat"), "w") as f:
        f.write("%d\n \n" % n)
        for mol in range(m):
            for i in range(ms[mol], ms[mol + 1]):
                f.write("%d %s %.17g %.17g %.17g %.17g\n" % (mol, names[ty[i]], xyz[i][0],
                                                             xyz[i][1], xyz[i][2], q[i]))
    b = sysdef["box"]
    with open(os.path.join(outdir, "input_top.dat"), "w") as f:
        f.write("TotNoOfBeads: %d\nTotNoOfMolec: %d\n" % (n, m))
        f.write("Box_Length_X: %.17g\nBox_Length_Y: %.17g\nBox_Length_Z: %.17g\n" % tuple(b))
        f.write("bonds:\n")
        for mol in range(m):
            for i in range(ms[mol], ms[mol + 1] - 1):
                f.write("%d %d %d\n" % (mol, i - ms[mol], i - ms[mol] + 1))
        f.write("-1\nangles:\n-1\ndiheds\n-1\n")
    mp = sysdef["move_prob"]
    lines = [
        ("s1_input_coordinate_file", "input_crd.dat"),
        ("s1_input_topology_file", "input_top.dat"),
        ("s1_prefix_of_output_file_names", prefix),
        ("s1_total_simulation_steps", steps),
        ("s1_equilibrium_steps", 0),
        ("s1_sampling_frequency", 100),
        ("s1_sampling_print_frequency", 1000),
        ("s1_trajectory_print_frequency", 100000),
        ("s1_PBC_dimensions", sysdef["npbc"]),
        ("s1_beta_(1/kBT)", sysdef["beta"]),
        ("s1_MC_move_size_in_unit_length", sysdef["move_size"]),
        ("s1_calc_chem_pot", 0),
        ("s1_number_of_surface_sites", sysdef.get("phantom", 0)),
        ("s1_number_of_surf_counterions", 0),
        ("s1_number_of_grafted_chains", sysdef.get("grafted", 0)),
        ("s1_number_of_grafted_counterion", 0),
        ("s1_prob_b_bead_translation", mp[0]),
        ("s1_prob_p_COM_translation", mp[1]),
        ("s1_prob_p_pivot", mp[2]),
        ("s1_prob_p_crankshaft", mp[3]),
        ("s1_prob_p_random_reptation", mp[4]),
        ("s1_vp_bin_resolution_in_ul", 10),
        ("s1_bead_size_virial_pressure", 1.0),
        ("s2_use_short_range_potential", sysdef["use_pair"]),
        ("s2_use_Ewald_potential", sysdef["use_ewald"]),
        ("s2_use_bond_potential", sysdef["use_bond"]),
        ("s2_use_rigid_bond", 1),
        ("s2_use_angle_potential", 0),
        ("s2_use_dihedral_potential", 0),
        ("s2_use_external_potential", sysdef["use_ext"]),
        ("s2_use_grand_canonical_MC_move", 1 if "gc" in sysdef else 0),
    ]
    if "gc" in sysdef:
        g = sysdef["gc"]
        lines += [("s3_chemical_potential_in_kT", repr(float(g["chem_pot"]))),
                  ("s3_deBroglie_prefactor", repr(float(g["debroglie"]))),
                  ("s3_GCMC_move_frequency", g["freq"]),
                  ("s3_chain_length", g["chain_len"]),
                  ("s3_bead_charge_in_e", int(g["bead_charge"])),
                  ("s3_bead_type", names[g.get("type", 0)]),
                  ("s3_no_of_CBMC_trials", g["trials"])]
    if sysdef["use_pair"]:
        lines.append(("s3_short_range_potential_type", "TruncatedLJ"))
        lines.append(("s3_LJ_cutoff", sysdef["lj_cutoff"]))
        for k, t in enumerate(sysdef["types"]):
            lines.append(("s3_bead_type_%d" % (k + 1), t["name"]))
            lines.append(("s3_sigma_for_bead_type_%d" % (k + 1), t["sigma"]))
            lines.append(("s3_epsilon_for_bead_type_%d" % (k + 1), t["eps"]))
        lines.append(("s3_end_flag", "end"))
    if sysdef["use_ewald"]:
        lines += [("s3_Ewald_potential_type", "Coul"), ("s3_Bjerrum_length", sysdef["lB"]),
                  ("s3_dielectric_constant", sysdef["dielectric"]),
                  ("s3_Ewald_alpha", repr(float(sysdef["alpha"]))),
                  ("s3_use_dipole_correction", sysdef["dipole"])]
    if sysdef["use_bond"]:
        lines += [("s3_bond_potential_type", "Spring"), ("s3_spring_k", sysdef["k_bond"]),
                  ("s3_spring_r0", sysdef["r0"])]
    lines.append(("s3_rigid_bond_length", sysdef["r0"]))
    if sysdef["use_ext"]:
        lines += [("s3_external_potential_type", "TruncatedLJWall"),
                  ("s3_wall_cutoff", sysdef["wall_cut"]),
                  ("s3_wall_sigma", sysdef["wall_sigma_w"]),
                  ("s3_wall_epsilon", sysdef["wall_eps_w"])]
        for k, t in enumerate(sysdef["types"]):
            lines.append(("s3_wall_bead_type_%d" % (k + 1), t["name"]))
            lines.append(("s3_wall_sigma_for_bead_type_%d" % (k + 1), t["wall_sigma"]))
            lines.append(("s3_wall_epsilon_for_bead_type_%d" % (k + 1), t["wall_eps"]))
        lines.append(("s3_wall_end_flag", "end"))
    with open(os.path.join(outdir, "run.in"), "w") as f:
        for k, v in lines:
            f.write("%-34s %s\n" % (k, v))
    return os.path.join(outdir, "run.in")
