"""Reads a reference input deck (run.in + crd + top) into the system
definition the oracle and the engine take.  Test helper: the positional
run.in grammar of simulation.cc:30-50, force_field.cc:65-195 and the
potentials' ReadParameters, and the crd / top readers of simulation.cc:582-646.
"""
import os

import numpy as np


def _pairs(path):
    toks = open(path).read().split()
    return [(toks[i], toks[i + 1]) for i in range(0, len(toks) - 1, 2)]


def deck_sysdef(run_in):
    """System definition (oracle/engine dict) and the move options of a deck
    whose crd/top files sit next to it."""
    kv = _pairs(run_in)
    it = iter(kv)

    def nxt():
        return next(it)[1]

    crd, top = nxt(), nxt()
    for _ in range(6):
        nxt()                                  # prefix, steps, eq, sampling, print freqs
    npbc = int(nxt())
    beta = float(nxt())
    move_size = float(nxt())
    nxt()                                      # calc_chem_pot
    phantom, coion, grafted, _gci = (int(nxt()) for _ in range(4))
    move_prob = [float(nxt()) for _ in range(5)]
    nxt(), nxt()                               # vp_el_res, vp_bead_size
    use_pair, use_ewald, use_bond, use_rigid, _ang, _dih, use_ext, use_gc = (
        int(nxt()) for _ in range(8))
    gc = None
    if use_gc:
        # chemical potential, de Broglie prefactor, move frequency, chain
        # length, bead charge, bead type, CBMC trials (force_field.cc:96-103)
        gc = dict(chem_pot=float(nxt()), debroglie=float(nxt()), freq=int(nxt()), chain_len=int(nxt()),
                  bead_charge=float(nxt()), type_name=nxt(), trials=int(nxt()), coion=coion)
    lj, wall = {}, {}
    lj_cutoff = 2.5
    lB = dielectric = alpha = 0.0
    dipole = 0
    k_bond, r0, rigid = 0.0, 1.0, 0.0
    wall_cut, wall_sigma_w, wall_eps_w = -1.0, 1.0, 1.0

    def table(dst):
        while True:
            sym = nxt()
            if sym == "end":
                return
            dst[sym] = (float(nxt()), float(nxt()))

    if use_pair:
        assert nxt() == "TruncatedLJ"
        lj_cutoff = float(nxt())
        table(lj)
    if use_ewald:
        assert nxt() == "Coul"
        lB, dielectric, alpha = float(nxt()), float(nxt()), float(nxt())
        dipole = int(nxt())
    if use_bond:
        assert nxt() == "Spring"
        k_bond, r0 = float(nxt()), float(nxt())
    if use_rigid:
        rigid = float(nxt())
    if use_ext:
        assert nxt() == "TruncatedLJWall"
        wall_cut, wall_sigma_w, wall_eps_w = float(nxt()), float(nxt()), float(nxt())
        table(wall)
    here = os.path.dirname(run_in)
    rows = [l.split() for l in open(os.path.join(here, crd)).read().splitlines()[2:] if l.strip()]
    tinfo = {r[0]: r[1] for r in (l.split() for l in open(os.path.join(here, top))) if len(r) > 1}
    box = [float(tinfo["Box_Length_%s:" % a]) for a in "XYZ"]
    n_mol = int(tinfo["TotNoOfMolec:"])
    names = []
    for s in list(lj) + list(wall) + [r[1] for r in rows]:
        if s not in names:
            names.append(s)
    types = [{"name": s, "sigma": lj.get(s, (0.0, 0.0))[0], "eps": lj.get(s, (0.0, 0.0))[1],
              "wall_sigma": wall.get(s, (0.0, 0.0))[0], "wall_eps": wall.get(s, (0.0, 0.0))[1]}
             for s in names]
    mol = np.array([int(r[0]) for r in rows])
    order = np.argsort(mol, kind="stable")       # beads appended to mols[mol_id] in file order
    xyz = np.array([[float(r[2]), float(r[3]), float(r[4])] for r in rows])[order]
    q = np.array([float(r[5]) for r in rows])[order]
    ty = np.array([names.index(r[1]) for r in rows], dtype=np.int32)[order]
    counts = np.bincount(mol, minlength=n_mol)
    ms = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    sd = dict(npbc=npbc, beta=beta, use_pair=use_pair, use_ewald=use_ewald, use_bond=use_bond,
              use_ext=use_ext, lj_cutoff=lj_cutoff, lB=lB, dielectric=dielectric, alpha=alpha,
              dipole=dipole, k_bond=k_bond, r0=r0, wall_cut=wall_cut, wall_sigma_w=wall_sigma_w,
              wall_eps_w=wall_eps_w, phantom=phantom, grafted=grafted, move_prob=move_prob,
              move_size=move_size, types=types, box=box, xyz=xyz, q=q, type=ty, mol_start=ms,
              name=os.path.basename(here))
    if gc is not None:
        gc["type"] = names.index(gc.pop("type_name"))
        sd["gc"] = gc
    bond_len = rigid if use_rigid else (r0 if use_bond else 0.0)
    vary = (not use_rigid) and bool(use_bond)
    return sd, dict(phantom=phantom, coion=coion, grafted=grafted, bond_len=bond_len,
                    vary_bond=vary)
