"""Host layer: the run.in / crd / top readers on the reference's own example
decks (CPU), the exported ABI, and the host Monte Carlo driver replaying the
reference move sequence against the oracle (GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from plum_amd import host as H
from plum_amd import systems

HERE = os.path.dirname(os.path.abspath(__file__))
EX = os.path.join(HERE, "golden", "examples")
ROOT = os.path.dirname(HERE)


def test_host_library_exports_every_declared_symbol():
    text = open(os.path.join(ROOT, "include", "plum_host.h")).read()
    decl = sorted(set(re.findall(r"\b(plum_[a-z_]+)\s*\(", text)))
    assert decl == sorted(H.HOST_SYMBOLS)
    lib = H.lib()
    for s in decl:
        assert hasattr(lib, s), s


def test_parse_reference_example_decks():
    s = H.parse_runin(os.path.join(EX, "bulk_nvt", "run.in"))
    assert (s.n_beads, s.n_mols, s.npbc) == (512, 272, 3)
    assert s.box[0] == pytest.approx(58.4803547642572994)
    assert (s.use_pair, s.use_ewald, s.use_bond, s.use_ext, s.use_gc) == (1, 1, 0, 0, 0)
    assert s.lj_cutoff == -1 and s.lB == 2.5 and s.alpha == 0.001 and s.rigid_bond == 2.5
    assert list(s.move_prob) == [0.95, 0.0, 0.05, 0.0, 0.0]
    assert s.n_types == 5  # P L R C Z
    c = H.parse_runin(os.path.join(EX, "confined_nvt", "run.in"))
    assert c.npbc == 2 and c.use_ext == 1 and c.dipole_correction == 1
    assert c.box[2] == 60.0
    g = H.parse_runin(os.path.join(EX, "bulk_muvt", "run.in"))
    assert g.use_gc == 1 and g.n_beads == 0


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_parse_errors_are_status_codes(tmp_path):
    # truncated deck: the reference would read garbage / exit(1)
    src = open(os.path.join(EX, "bulk_nvt", "run.in")).read().splitlines()
    for f in ("input_top.dat", "input_crd.dat"):
        (tmp_path / f).write_text(open(os.path.join(EX, "bulk_nvt", f)).read())
    p = _write(tmp_path, "run.in", "\n".join(src[:30]) + "\n")
    with pytest.raises(H.HostError) as e:
        H.parse_runin(p)
    assert "run.in" in str(e.value)
    # molecule id out of range in the coordinate file (simulation.cc:594-598)
    crd = open(os.path.join(EX, "bulk_nvt", "input_crd.dat")).read().splitlines()
    crd[2] = "272 " + crd[2].split(" ", 1)[1]
    (tmp_path / "input_crd.dat").write_text("\n".join(crd) + "\n")
    p = _write(tmp_path, "run.in", "\n".join(src) + "\n")
    with pytest.raises(H.HostError) as e:
        H.parse_runin(p)
    assert "out of bounds" in str(e.value)


def test_generated_decks_round_trip(tmp_path):
    sd = systems.config("slab_small")
    path = systems.write_reference_inputs(sd, str(tmp_path))
    s = H.parse_runin(path)
    assert s.n_beads == len(sd["q"]) and s.n_mols == len(sd["mol_start"]) - 1
    assert s.npbc == 2 and s.use_ext == 1 and s.dipole_correction == 1
    assert s.alpha == pytest.approx(sd["alpha"], rel=1e-15)


@pytest.mark.gpu
@pytest.mark.parametrize("name,steps", [("small", 120), ("slab_small", 120)])
def test_host_driver_replays_reference_sequence(tmp_path, name, steps):
    """Same run.in, same seed: the host driver (GPU energies) and the oracle
    (CPU reference algorithm) take the same accept/reject decisions and end
    in the same configuration."""
    from oracle import oracle as O
    sd = systems.config(name)
    path = systems.write_reference_inputs(sd, str(tmp_path))
    sim = H.HostSimulation(path, 77)
    att, acc = sim.run(steps)
    orc = O.Oracle(sd)
    rng = O.MT19937(77)
    oatt, oacc = [0] * 5, [0] * 5
    for _ in range(steps):
        inf = O.mc_step(orc, rng, sd["move_prob"], sd["move_size"], sd.get("phantom", 0),
                        sd.get("grafted", 0), sd["r0"], True)
        if inf.attempted:
            oatt[inf.move_type] += 1
            oacc[inf.move_type] += inf.accepted
    print(name, "host", att, acc, "oracle", oatt, oacc, flush=True)
    assert att == oatt and acc == oacc
    pos = sim.positions(len(sd["q"]))
    assert np.max(np.abs(pos - orc.positions())) <= 1e-9
    run, full = sim.energies()
    ref = orc.energy_full()
    for k in ref:
        assert abs(full[k] - ref[k]) <= 1e-10 * max(abs(ref[k]), 1e-2), k


@pytest.mark.gpu
@pytest.mark.parametrize("mu,steps", [(-45.0, 400), (-60.0, 240)])
def test_host_grand_canonical_run_replays_reference_sequence(tmp_path, mu, steps):
    """Grand-canonical chain exchange with configurational bias: the host
    driver (trial energies of each growth step in one GPU pass) and the
    oracle take the same insertions, deletions and trial moves on the same
    seed, end with the same molecules at the same positions, and their Widom
    test insertions agree.  At the lower chemical potential the system
    empties and refills."""
    from oracle import oracle as O
    sd = systems.config("gc_small")
    sd["gc"] = dict(sd["gc"], chem_pot=mu)
    path = systems.write_reference_inputs(sd, str(tmp_path))
    sim = H.HostSimulation(path, 31)
    att, acc = sim.run(steps)
    gatt, gacc, nb, nm = sim.gc_stats()
    orc = O.GCSim(sd, 31)
    orc.run(steps)
    print("mu %.0f host gc %s %s trial %s %s | oracle gc %s %s trial %s %s | beads %d" % (
        mu, gatt, gacc, att, acc, orc.gc_attempted, orc.gc_accepted, orc.attempted,
        orc.accepted, nb), flush=True)
    assert gatt == orc.gc_attempted and gacc == orc.gc_accepted
    assert att == orc.attempted and acc == orc.accepted
    assert nb == len(orc.sd["q"]) and nm == len(orc.sd["mol_start"]) - 1
    if nb:
        assert np.max(np.abs(sim.positions(nb) - orc.orc.positions())) <= 1e-9
    w_host, w_orc = sim.widom(), orc.widom()
    assert abs(w_host - w_orc) <= 1e-9 * max(abs(w_orc), 1e-300), (w_host, w_orc)
    run, full = sim.energies()
    for k in full:
        assert abs(run[k] - full[k]) <= 1e-9 * max(1.0, abs(full[k])), k
    orc.close()
    sim.close()


@pytest.mark.gpu
def test_host_grand_canonical_device_run_bookkeeping(tmp_path):
    """Exchanges between batches of device-loop moves: the running totals
    stay the from-scratch energy, and the host molecules follow the device
    coordinates."""
    sd = systems.config("gc_small")
    sd["gc"] = dict(sd["gc"], chem_pot=-45.0)
    path = systems.write_reference_inputs(sd, str(tmp_path))
    sim = H.HostSimulation(path, 12)
    for _ in range(4):
        att, acc, sec = sim.run_device(300)
        gatt, gacc, nb, nm = sim.gc_stats()
        run, full = sim.energies()
        for k in full:
            assert abs(run[k] - full[k]) <= 1e-9 * max(1.0, abs(full[k])), k
    print("device run: gc %s %s, moves %s %s, beads %d" % (gatt, gacc, att, acc, nb), flush=True)
    assert sum(att) > 100 and sum(gatt) > 50
    sim.close()


def _deck_sysdef(path):
    """System definition of a reference run.in whose crd/top hold no beads
    (the grand-canonical example), read by parameter name."""
    rows = [line.split() for line in open(path) if line.strip()]
    d = {r[0]: r[1] for r in rows}
    types, k = [], 1
    while "s3_bead_type_%d" % k in d:
        types.append({"name": d["s3_bead_type_%d" % k],
                      "sigma": float(d["s3_sigma_for_bead_type_%d" % k]),
                      "eps": float(d["s3_epsilon_for_bead_type_%d" % k]),
                      "wall_sigma": 0.0, "wall_eps": 0.0})
        k += 1
    mp = [float(d[x]) for x in ("s1_prob_b_bead_translation", "s1_prob_p_COM_translation",
                                "s1_prob_p_pivot", "s1_prob_p_crankshaft",
                                "s1_prob_p_random_reptation")]
    top = {r[0]: r[1] for r in (line.split() for line in
                                open(os.path.join(os.path.dirname(path), "input_top.dat")))
           if len(r) > 1}
    box = [float(top["Box_Length_%s:" % a]) for a in "XYZ"]
    return dict(npbc=int(d["s1_PBC_dimensions"]), beta=float(d["s1_beta_(1/kBT)"]),
                use_pair=int(d["s2_use_short_range_potential"]),
                use_ewald=int(d["s2_use_Ewald_potential"]), use_bond=int(d["s2_use_bond_potential"]),
                use_ext=0, lj_cutoff=float(d["s3_LJ_cutoff"]), lB=float(d["s3_Bjerrum_length"]),
                dielectric=float(d["s3_dielectric_constant"]), alpha=float(d["s3_Ewald_alpha"]),
                dipole=int(d["s3_use_dipole_correction"]), k_bond=0.0,
                r0=float(d["s3_rigid_bond_length"]), wall_cut=-1.0, wall_sigma_w=1.0,
                wall_eps_w=1.0, phantom=0, grafted=0, move_prob=mp,
                move_size=float(d["s1_MC_move_size_in_unit_length"]), types=types, box=box,
                xyz=np.zeros((0, 3)), q=np.zeros(0), type=np.zeros(0, dtype=np.int32),
                mol_start=np.zeros(1, dtype=np.int32),
                gc=dict(chain_len=int(d["s3_chain_length"]),
                        bead_charge=float(d["s3_bead_charge_in_e"]),
                        trials=int(d["s3_no_of_CBMC_trials"]), freq=int(d["s3_GCMC_move_frequency"]),
                        chem_pot=float(d["s3_chemical_potential_in_kT"]),
                        debroglie=float(d["s3_deBroglie_prefactor"]), type=0))


@pytest.mark.gpu
def test_reference_grand_canonical_example_replays_on_the_oracle():
    """The reference's own grand-canonical deck (empty box of 100, 8-bead chains of
    charge -1 with counterions, 30 trials, mu = -90, rigid bonds, WCA, l_B 2.5,
    alpha 0.001 so 125 real-space images): host driver and oracle fill the box
    with the same exchanges and trial moves."""
    from oracle import oracle as O
    path = os.path.join(EX, "bulk_muvt", "run.in")
    sd = _deck_sysdef(path)
    steps = 800
    sim = H.HostSimulation(path, 7)
    att, acc = sim.run(steps)
    gatt, gacc, nb, nm = sim.gc_stats()
    orc = O.GCSim(sd, 7)
    orc.run(steps)
    print("example deck: host gc %s %s trial %s %s | oracle gc %s %s trial %s %s | beads %d" % (
        gatt, gacc, att, acc, orc.gc_attempted, orc.gc_accepted, orc.attempted, orc.accepted,
        nb), flush=True)
    assert gatt == orc.gc_attempted and gacc == orc.gc_accepted
    assert att == orc.attempted and acc == orc.accepted
    assert nb == len(orc.sd["q"]) and sum(gacc) > 0
    if nb:
        assert np.max(np.abs(sim.positions(nb) - orc.orc.positions())) <= 1e-9
    orc.close()
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("mode,log_theta", [(1, 0.0), (2, -16.0), (2, -6.0)])
def test_two_stage_exchange_replays_on_the_oracle(mode, log_theta):
    """The reference's grand-canonical deck with the two-stage exchange
    (plum_sim_gc_prescreen mode 2: first stage on the first bead's exact
    weight times theta, growth and second stage only past it) against the
    oracle's restatement (orc_gc_insertion2 / orc_gc_deletion2) on the same
    seed: the same first-stage outcomes, exchanges, trial moves and final
    configuration.  Mode 1 (the reference's exchange, recording ln(W / w_1)
    of each growth) must leave the reference's chain untouched."""
    from oracle import oracle as O
    path = os.path.join(EX, "bulk_muvt", "run.in")
    sd = _deck_sysdef(path)
    steps = 800
    sim = H.HostSimulation(path, 7)
    sim.gc_prescreen(mode, log_theta)
    att, acc = sim.run(steps)
    gatt, gacc, nb, nm = sim.gc_stats()
    rej, pas, mean_rest, _, n_rest = sim.gc_prescreen_stats()
    orc = O.GCSim(sd, 7, log_theta=log_theta if mode == 2 else None)
    orc.run(steps)
    print("mode %d log theta %.1f: host gc %s %s, first stage rejected %s passed %s, mean ln(W/w1) %.3f "
          "over %d | oracle gc %s %s, first stage rejected %s passed %s" % (
              mode, log_theta, gatt, gacc, rej, pas, mean_rest, n_rest, orc.gc_attempted,
              orc.gc_accepted, orc.s1_rejected, orc.s1_passed), flush=True)
    assert gatt == orc.gc_attempted and gacc == orc.gc_accepted
    assert att == orc.attempted and acc == orc.accepted
    assert nb == len(orc.sd["q"])
    if mode == 2:
        assert rej == orc.s1_rejected and pas == orc.s1_passed and sum(pas) > 0
    else:
        assert rej == [0, 0] and pas == [0, 0] and n_rest > 0
    if nb:
        assert np.max(np.abs(sim.positions(nb) - orc.orc.positions())) <= 1e-9
    orc.close()
    sim.close()


@pytest.mark.gpu
def test_reference_confined_grand_canonical_example_replays_on_the_oracle():
    """The reference's own confined grand-canonical deck (18 surface sites and
    their counterions between walls, slab Ewald with the dipole correction,
    8-bead chains of charge -1 with counterions, 30 trials, mu = -90): every
    growth step's trial energies carry the slab dipole term (DipoleEDiff,
    potential_ewald_coul.cc:473-530); the host driver with GPU growths and
    the oracle take the same exchanges and trial moves and end in the same
    configuration."""
    from deck import deck_sysdef
    from oracle import oracle as O
    path = os.path.join(EX, "confined_muvt", "run.in")
    sd, mv = deck_sysdef(path)
    sd["r0"] = mv["bond_len"]              # rigid bonds: the reptation and growth length
    steps = 800
    sim = H.HostSimulation(path, 11)
    att, acc = sim.run(steps)
    gatt, gacc, nb, nm = sim.gc_stats()
    orc = O.GCSim(sd, 11)
    orc.run(steps)
    print("confined example deck: host gc %s %s trial %s %s | oracle gc %s %s trial %s %s | beads %d" % (
        gatt, gacc, att, acc, orc.gc_attempted, orc.gc_accepted, orc.attempted, orc.accepted,
        nb), flush=True)
    assert gatt == orc.gc_attempted and gacc == orc.gc_accepted
    assert att == orc.attempted and acc == orc.accepted
    assert nb == len(orc.sd["q"]) and sum(gacc) > 0
    assert np.max(np.abs(sim.positions(nb) - orc.orc.positions())) <= 1e-9
    run, full = sim.energies()
    ref = orc.orc.energy_full()
    assert abs(full["dipole"] - ref["dipole"]) <= 1e-10 * max(abs(ref["dipole"]), 1e-2)
    orc.close()
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["small", "slab_small"])
def test_last_state_files_resume_the_run(tmp_path, name):
    """PrintLastCrd / PrintLastTop: the written configuration, read back by a
    run.in that names it, is the same system up to whole-box shifts of
    molecules, with the same energy; the density profile matches its
    definition (simulation.cc:1029-1045) on the same coordinates."""
    sd = systems.config(name)
    path = systems.write_reference_inputs(sd, str(tmp_path))
    sim = H.HostSimulation(path, 9)
    sim.run(400)
    n = len(sd["q"])
    x0 = sim.positions(n)
    _, e0 = sim.energies()
    rho = sim.density_profile(1.0)
    sim.write_last(str(tmp_path / "out"))
    lines = open(path).read().splitlines()
    lines[0] = "s1_input_coordinate_file out_lastcrd.dat"
    lines[1] = "s1_input_topology_file out_lasttop.dat"
    resume = tmp_path / "resume.in"
    resume.write_text("\n".join(lines) + "\n")
    sim2 = H.HostSimulation(str(resume), 9)
    x1 = sim2.positions(n)
    _, e1 = sim2.energies()
    box = np.asarray(sd["box"])
    ms = sd["mol_start"]
    for m in range(len(ms) - 1):
        d = (x1[ms[m]:ms[m + 1]] - x0[ms[m]:ms[m + 1]]) / box
        k = np.round(d)
        assert np.abs(d - k).max() < 1e-12 and (k == k[0]).all(), m
    for t in e0:
        assert abs(e1[t] - e0[t]) <= 1e-9 * max(1.0, abs(e0[t])), (t, e0[t], e1[t])
    from test_gpu_ensemble import _rho_z
    ref = _rho_z(x0, sd["q"], ms, sd["box"], 1.0, sd.get("phantom", 0))
    assert np.abs(rho - ref).max() <= 1e-12
    sim.close()
    sim2.close()
