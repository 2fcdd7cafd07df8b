"""Ensemble averages of the device loop against the reference's own Markov
chain.  The reference chain is the host driver (plum_sim_run): the
reference's move sequence on std::mt19937, decision for decision the
oracle's (tests/test_host.py).  The device loop (plum_mc_run: Philox
streams, speculative batches) must sample the same ensemble: energy, radius
of gyration, end-to-end distance and the density profiles rho(z) of chain
beads and counterions agree within 2 sigma (block averages), both chains
starting from the same configuration equilibrated by the device loop, each
mean's error taken from the spread of independent chains."""
import numpy as np
import pytest

from plum_amd import engine as E
from plum_amd import host as H
from plum_amd import systems

pytestmark = pytest.mark.gpu


def _chains(xyz, ms):
    rg, ree = [], []
    for m in range(len(ms) - 1):
        a, b = ms[m], ms[m + 1]
        if b - a < 2:
            continue
        c = xyz[a:b]
        n = b - a
        d = c[:, None, :] - c[None, :, :]
        rg.append((d * d).sum() / 2 / (n * n))      # simulation.cc:406-434
        e = c[-1] - c[0]
        ree.append(e @ e)                           # simulation.cc:483-501
    return float(np.mean(rg)), float(np.mean(ree))


def _rho_z(xyz, q, ms, box, res, phantom):
    """Simulation::CalcRhoZ (simulation.cc:1029-1045): chain beads, cations,
    anions per z bin (bead counts / (Lx Ly res))."""
    nb = int(np.floor(box[2] / res))
    out = np.zeros((nb, 3))
    for m in range(phantom, len(ms) - 1):
        a, b = ms[m], ms[m + 1]
        sp = 0 if b - a > 1 else (1 if q[a] >= 0 else 2)
        z = xyz[a:b, 2] - box[2] * np.floor(xyz[a:b, 2] / box[2])
        bins = np.floor(z / res).astype(int)
        ok = (bins >= 0) & (bins < nb)
        np.add.at(out[:, sp], bins[ok], 1.0 / (box[0] * box[1] * res))
    return out


def _run_chains(run_one, n_chains):
    """Per-chain time averages of (energy, Rg^2, Ree^2) and rho(z)."""
    scal, prof = [], []
    for c in range(n_chains):
        o, p = run_one(c)
        scal.append(np.mean(o, 0))
        prof.append(np.mean(p, 0))
    return np.array(scal), np.array(prof)


def test_device_loop_samples_the_reference_chain_ensemble(tmp_path):
    """Independent chains on each side (the slab with walls, grafted chains
    and the lagged slab dipole, whose conformations decorrelate slowly): the
    error of each mean is the chain-to-chain spread."""
    name = "slab_small"
    sd0 = systems.config(name)
    mc = dict(move_prob=sd0["move_prob"], move_size=sd0["move_size"], bond_len=sd0["r0"],
              phantom=sd0.get("phantom", 0), grafted=sd0.get("grafted", 0))
    eq = E.Engine(sd0)
    eq.mc_run(40000, seed=11, **mc)
    sd = dict(sd0, xyz=eq.positions())
    eq.close()
    box, ms, q = sd["box"], sd["mol_start"], sd["q"]
    n, ph = len(q), sd.get("phantom", 0)
    n_chains, samples, chunk, skip = 6, 400, 300, 100
    path = systems.write_reference_inputs(sd, str(tmp_path))

    def host_chain(c):
        sim = H.HostSimulation(path, 100 + c)
        o, p = [], []
        for i in range(samples):
            sim.run(chunk)
            if i < skip:
                continue
            x = sim.positions(n)
            run, _ = sim.energies()
            o.append((sum(run[k] for k in E.TERMS), *_chains(x, ms)))
            p.append(_rho_z(x, q, ms, box, 1.0, ph))
        sim.close()
        return o, p

    def device_chain(c):
        eng = E.Engine(sd)
        o, p = [], []
        for i in range(samples):
            eng.mc_run(chunk, seed=(c + 1) * 1000003 + i, **mc)
            if i < skip:
                continue
            x = eng.positions()
            t = eng.totals()
            o.append((sum(t[k] for k in E.TERMS), *_chains(x, ms)))
            p.append(_rho_z(x, q, ms, box, 1.0, ph))
        eng.close()
        return o, p

    hs, hp = _run_chains(host_chain, n_chains)
    gs, gp = _run_chains(device_chain, n_chains)
    err = lambda a: a.std(0, ddof=1) / np.sqrt(len(a))
    z = np.abs(hs.mean(0) - gs.mean(0)) / np.hypot(err(hs), err(gs))
    for j, lab in enumerate(("energy", "Rg^2", "Ree^2")):
        print("%s: reference chain %.4f +- %.4f, device loop %.4f +- %.4f, |z| %.2f" % (
            lab, hs.mean(0)[j], err(hs)[j], gs.mean(0)[j], err(gs)[j], z[j]), flush=True)
    e = np.hypot(err(hp), err(gp))
    live = e > 0
    zp = np.abs(hp.mean(0) - gp.mean(0))[live] / e[live]
    chi2 = float((zp ** 2).mean())
    print("rho(z): %d populated bins, mean z^2 %.2f, bins beyond 2 sigma %d (expected %.1f)" % (
        live.sum(), chi2, int((zp > 2).sum()), 0.0455 * live.sum()), flush=True)
    assert z.max() < 2.0
    assert chi2 < 2.0 and (zp > 2).sum() <= max(2, int(np.ceil(0.15 * live.sum())))
