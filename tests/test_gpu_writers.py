"""The reference's Run loop with its sampling and output files
(Simulation::Run, simulation.cc:216-239: Sample :723-798, PrintStatHeader /
PrintStat :685-867, PrintTraj :869-933, PrintLastCrd / PrintLastTop
:935-1027, PrintLastRhoZ :1047-1064) through plum_sim_run_output, on the
reference's own bulk NVT deck.

The oracle replays the same steps from the same seed (the writers draw no
random numbers), and every sampled quantity is recomputed from the oracle's
energies and coordinates at the sample steps: the stat file's running
averages of the energies, radii of gyration and end-to-end distance, the
trajectory frames and the rho(z) histogram.  The stat columns are compared
at the precision the reference prints them with (7 significant digits for
the energies, 4 after).
"""
import os

import numpy as np
import pytest

from deck import deck_sysdef
from oracle import oracle as O
from plum_amd import host as H

pytestmark = pytest.mark.gpu
EX = os.path.join(os.path.dirname(__file__), "golden", "examples")
HEADER = ("#Step <PairEne> <EwaldEne> <TotalEnergy> <P> <|Rg|> <|Rgx|> <|Rgy|> <|Rgz|> <ete> "
          "beadtrans comtrans pivot crankshaft reptate")


def _chains(sd):
    ms = np.asarray(sd["mol_start"])
    return [(ms[i], ms[i + 1]) for i in range(len(ms) - 1) if ms[i + 1] - ms[i] > 1]


def _rg(xyz, chains):
    tot, ax = 0.0, np.zeros(3)
    for a, b in chains:
        p = xyz[a:b]
        d = p[:, None, :] - p[None, :, :]
        iu = np.triu_indices(b - a, 1)
        dd = d[iu[0], iu[1]]
        tot += np.sum(dd * dd) / (b - a) ** 2
        ax += np.sum(dd * dd, axis=0) / (b - a) ** 2
    n = len(chains)
    return tot / n, ax / n


def _ete(xyz, chains):
    return np.mean([np.sum((xyz[a] - xyz[b - 1]) ** 2) for a, b in chains])


def _close(text, ref, digits):
    got = float(text)
    return abs(got - ref) <= 10.0 ** (1 - digits) * max(abs(ref), 1e-300)


def test_stat_traj_rhoz_writers_follow_the_reference_run_loop(tmp_path):
    path = os.path.join(EX, "bulk_nvt", "run.in")
    sd, mv = deck_sysdef(path)
    steps, seed = 300, 2017
    prefix = str(tmp_path / "out")
    sim = H.HostSimulation(path, seed)
    sim.run_output(steps, prefix)
    # the oracle's replay, sampled where the deck samples (eq 100, every 100)
    orc = O.Oracle(sd)
    rng = O.MT19937(seed)
    samples = {}
    for step in range(1, steps + 1):
        O.mc_step(orc, rng, sd["move_prob"], sd["move_size"], mv["phantom"], mv["grafted"],
                  mv["bond_len"], mv["vary_bond"])
        if step > 100 and step % 100 == 0:
            samples[step] = (orc.running(), orc.positions().copy())
    chains = _chains(sd)
    box = np.asarray(sd["box"])

    stat = open(prefix + "_stat.dat").read().splitlines()
    assert stat[0] == HEADER
    rows = [r.split() for r in stat[1:]]
    assert [r[0] for r in rows] == ["100", "200", "300"]
    assert all(len(r) == len(HEADER.split()) for r in rows)
    # step 100 is not past equilibration: no sample yet, 0/0 averages
    assert rows[0][1:4] == ["-nan", "-nan", "-nan"] and rows[0][4] == "nan"
    for row, upto in ((rows[1], (200,)), (rows[2], (200, 300))):
        pair = np.mean([samples[s][0]["lj"] for s in upto])
        ewald = np.mean([samples[s][0]["real"] + samples[s][0]["kspace"] + samples[s][0]["self"]
                         + samples[s][0]["dipole"] for s in upto])
        assert _close(row[1], pair, 7) and _close(row[2], ewald, 7), (row, pair, ewald)
        assert _close(row[3], pair + ewald, 6)
        assert row[4] == "nan"
        rg = [_rg(samples[s][1], chains) for s in upto]
        rg_tot = np.sqrt(np.mean([g[0] for g in rg]))
        rg_ax = np.sqrt(np.mean([g[1] for g in rg], axis=0))
        ete = np.sqrt(np.mean([_ete(samples[s][1], chains) for s in upto]))
        assert _close(row[5], rg_tot, 4) and _close(row[9], ete, 4), (row, rg_tot, ete)
        for k in range(3):
            assert _close(row[6 + k], rg_ax[k], 4)

    # trajectory: one frame of the chains and one of the cations per sample step
    nchain = sum(b - a for a, b in chains)
    lines = open(prefix + "_traj_p.xyz").read().splitlines()
    assert len(lines) == 2 * (nchain + 2)
    assert lines[0] == str(nchain) and lines[1] == "STEP: 200"
    assert lines[nchain + 2] == str(nchain) and lines[nchain + 3] == "STEP: 300"
    frame = np.array([[float(v) for v in ln.split()[1:]] for ln in lines[nchain + 4:]])
    ref = np.concatenate([samples[300][1][a:b] for a, b in chains])
    # each molecule shifted by whole boxes so its centre lies in the box
    shift = (ref - frame) / box
    assert np.allclose(shift, np.round(shift), atol=1e-4)
    assert np.max(np.abs(frame - ref + np.round(shift) * box)) <= 1e-3
    ions = open(prefix + "_traj_c.xyz").read().splitlines()
    ncat = int(np.sum((np.diff(sd["mol_start"]) == 1) & (np.asarray(sd["q"])[np.asarray(sd["mol_start"][:-1])] > 0)))
    assert ions[0] == str(ncat) and len(ions) == 2 * (ncat + 2)

    # rho(z): bins of 0.02 over the box, the mean over the two samples;
    # integrated over z it gives the bead count of each species
    for suffix in ("", "2"):
        rho = np.loadtxt(prefix + "_lastrhoZ.dat" + suffix)
        nb = int(np.ceil(box[2] / 0.02))
        assert rho.shape == (nb, 4)
        assert np.allclose(rho[:, 0], (np.arange(nb) + 0.5) * 0.02)
        counts = rho[:, 1:].sum(axis=0) * box[0] * box[1] * 0.02
        q = np.asarray(sd["q"])
        ms = np.asarray(sd["mol_start"])
        single = np.zeros(len(q), bool)
        for i in range(len(ms) - 1):
            single[ms[i]:ms[i + 1]] = ms[i + 1] - ms[i] == 1
        ref_counts = [np.sum(~single), np.sum(single & (q >= 0)), np.sum(single & (q < 0))]
        assert np.allclose(counts, ref_counts, rtol=1e-5), (counts, ref_counts)
    assert os.path.exists(prefix + "_lastcrd.dat") and os.path.exists(prefix + "_lastcrd.dat2")
    assert open(prefix + "_lasttop.dat").readline().startswith("TotNoOfBeads: %d" % len(sd["q"]))
    orc.close()
    sim.close()


def test_device_batches_between_output_steps_keep_the_files_consistent(tmp_path):
    """With the translational moves on the device loop (flushed before every
    sampling or output step) the files have the same layout, the stat rows
    average the energies the engine reports at the sample steps, and the
    last configuration written is the engine's."""
    path = os.path.join(EX, "bulk_nvt", "run.in")
    sd, _ = deck_sysdef(path)
    prefix = str(tmp_path / "dev")
    sim = H.HostSimulation(path, 11)
    sim.run_output(200, prefix, device=True)
    run, full = sim.energies()
    stat = open(prefix + "_stat.dat").read().splitlines()
    rows = [r.split() for r in stat[1:]]
    assert [r[0] for r in rows] == ["100", "200"]
    assert _close(rows[1][1], full["lj"], 6)
    crd = open(prefix + "_lastcrd.dat").read().splitlines()
    assert int(crd[0]) == len(sd["q"])
    xyz = np.array([[float(v) for v in ln.split()[2:5]] for ln in crd[2:]])
    pos = sim.positions(len(sd["q"]))
    box = np.asarray(sd["box"])
    shift = (pos - xyz) / box
    assert np.allclose(shift, np.round(shift), atol=1e-9)
    sim.close()


def test_slit_pressure_columns_on_the_reference_confined_deck(tmp_path):
    """With walls the stat file carries the six wall-force slit pressure
    averages (ForceField::CalcPressureForceLJELSlit every sample_freq * 10
    steps, pressure.cc:404-487; GetPressure :490-500): ion,
    chain and wall-wall LJ, then the same for the Ewald forces (the header's
    <Pzz_LJ_ion> .. <Pzz_EL_wal>), each force sum over the area (the
    wall-wall ones from the first sample, the others averaged over the
    samples).  The oracle replays the reference's chain from
    the same seed and sums the same forces at the sample steps (1000, 2000);
    the columns agree to the 6 significant digits they are printed with."""
    path = os.path.join(EX, "confined_nvt", "run.in")
    sd, mv = deck_sysdef(path)
    steps, seed = 2000, 2017
    prefix = str(tmp_path / "out")
    sim = H.HostSimulation(path, seed)
    sim.run_output(steps, prefix)
    orc = O.Oracle(sd)
    rng = O.MT19937(seed)
    fs = []
    for step in range(1, steps + 1):
        O.mc_step(orc, rng, sd["move_prob"], sd["move_size"], mv["phantom"], mv["grafted"],
                  mv["bond_len"], mv["vary_bond"])
        if step > 100 and step % 1000 == 0:
            fs.append(orc.slit_forces(mv["phantom"]))
    area = sd["box"][0] * sd["box"][1]
    fs = np.array(fs)
    want = [fs[:, 0].mean() / area, fs[:, 1].mean() / area, fs[0, 2] / area,
            fs[:, 3].mean() / area, fs[:, 4].mean() / area, fs[0, 5] / area]
    lines = open(prefix + "_stat.dat").read().splitlines()
    head = lines[0].split()
    rows = [r.split() for r in lines[1:]]
    last = [r for r in rows if r[0] == str(steps)][0]
    assert len(last) == len(head)
    c = head.index("<Pzz_LJ_ion>")
    assert head[c:c + 6] == ["<Pzz_LJ_ion>", "<Pzz_LJ_pol>", "<Pzz_LJ_wal>", "<Pzz_EL_ion>", "<Pzz_EL_pol>",
                             "<Pzz_EL_wal>"]
    got = last[c:c + 6]
    print("stat", got, "oracle", want, flush=True)
    assert any(abs(w) > 0 for w in want)
    for text, ref in zip(got, want):
        assert _close(text, ref, 6), (got, want)
    sim.close()
