"""Delayed acceptance with the learned surrogate on the GPU: the inference
(BF16 matrix cores) and training (FP32) kernels against the float64
reference in oracle/surrogate.py, the descriptors against the reference, the
ensemble sampled with delayed acceptance against the plain chain, and what
training the network gains."""
import numpy as np
import pytest

from oracle import surrogate as S
from plum_amd import engine as E
from plum_amd import systems

pytestmark = pytest.mark.gpu


def _mc(sd, seed):
    return dict(move_prob=sd["move_prob"], move_size=sd["move_size"], seed=seed,
                phantom=sd.get("phantom", 0), grafted=sd.get("grafted", 0), bond_len=1.0)


def _engine(name="small", **kw):
    sd = systems.config(name)
    eng = E.Engine(sd)
    eng.surrogate_enable(**kw)
    return sd, eng


def test_inference_and_training_forward_match_reference():
    sd, eng = _engine(train_every=0, seed=4)
    w = eng.surrogate_weights()
    rng = np.random.default_rng(1)
    x = rng.normal(0, 1.0, (300, S.DESCR)).astype(np.float32)
    f32 = eng.surrogate_forward(x, bf16=False)
    ref = S.forward(w, x)
    scale = np.abs(ref).max()
    err32 = np.abs(f32 - ref).max() / scale
    fbf = eng.surrogate_forward(x, bf16=True)
    refbf = S.forward(w, x, round_bf16=True)
    errbf = np.abs(fbf - refbf).max() / scale
    errbf_full = np.abs(fbf - ref).max() / scale
    print("forward: fp32 rel %.2e, bf16 vs bf16-rounded reference %.2e, vs fp64 %.2e" % (
        err32, errbf, errbf_full), flush=True)
    assert err32 < 1e-5
    # matrix-core layout and BF16 rounding reproduced: only activations that
    # round differently at a BF16 tie boundary can differ
    assert errbf < 2e-3
    assert errbf_full < 5e-2
    eng.close()


def test_training_gradient_matches_reference():
    sd, eng = _engine(train_every=0, seed=6)
    maxm = eng.max_mol_len()
    w = eng.surrogate_weights()
    rng = np.random.default_rng(2)
    samples = 5
    x = rng.normal(0, 1.0, (samples * 2 * maxm, S.DESCR)).astype(np.float32)
    n = rng.integers(1, maxm + 1, samples).astype(np.int32)
    target = rng.normal(0, 3.0, samples)
    loss, grad = eng.surrogate_gradient(x, target, n)
    rloss, rgrad = S.loss_and_grad(w, x, target, n, maxm, clip=20.0)
    gerr = np.abs(grad - rgrad).max() / np.abs(rgrad).max()
    print("gradient: loss %.6e vs %.6e, grad max rel %.2e" % (loss, rloss, gerr), flush=True)
    assert abs(loss - rloss) <= 1e-4 * rloss
    assert gerr < 1e-4
    eng.close()


@pytest.mark.parametrize("name", ["small", "slab_small"])
def test_descriptors_match_reference(name):
    sd, eng = _engine(name, screen=1, fill=0, train_every=0, seed=2)
    before = eng.positions()
    eng.mc_run(1, **_mc(sd, 77))
    rows, pos, n = eng.surrogate_debug_slot(0)
    maxm = eng.max_mol_len()
    assert n > 0
    # the moved beads: their old positions are the pre-batch positions
    lo = int(np.flatnonzero((before == pos[0, :3]).all(1))[0])
    excl = (lo, lo + n - 1)
    box = sd["box"]
    worst = 0.0
    for m in range(n):
        for on in (0, 1):
            p = pos[on * maxm + m]
            ref = S.descriptor(p[:3], p[3], before, sd["q"], excl, box, sd["npbc"], 4.0)
            got = rows[on * maxm + m]
            worst = max(worst, np.abs(got - ref).max() / max(1.0, np.abs(ref).max()))
    print("%s descriptors (%d beads): worst rel %.2e" % (name, n, worst), flush=True)
    assert worst < 2e-4
    eng.close()


def _observables(eng, sd):
    xyz = eng.positions()
    ms = sd["mol_start"]
    rg = []
    for m in range(len(ms) - 1):
        a, b = ms[m], ms[m + 1]
        if b - a > 1:
            c = xyz[a:b] - xyz[a:b].mean(0)
            rg.append((c * c).sum(1).mean())
    t = eng.totals()
    return sum(t[k] for k in E.TERMS if k != "dipole"), float(np.mean(rg))


def _block_stats(v, blocks=10):
    v = np.asarray(v)
    b = v[: len(v) // blocks * blocks].reshape(blocks, -1).mean(1)
    return b.mean(), b.std(ddof=1) / np.sqrt(blocks)


def test_delayed_acceptance_samples_the_plain_ensemble():
    """Energy and radius of gyration of delayed-acceptance chains agree with
    plain chains within 2 sigma, sigma from the spread of twelve independent
    chains on each side (with six the spread itself is too uncertain for a
    2 sigma bound), all starting from one configuration equilibrated by a long
    plain run.  The network is trained online first and then held fixed for
    the measured chains; every delayed-acceptance chain keeps its running
    totals equal to its from-scratch energy."""
    name = "small"
    sd0 = systems.config(name)
    eq = E.Engine(sd0)
    eq.mc_run(30000, **_mc(sd0, 41))
    sd = dict(sd0, xyz=eq.positions())
    eq.close()
    tr = E.Engine(sd)
    tr.surrogate_enable(screen=16, fill=2, train_every=8, train_samples=32, seed=3)
    tr.mc_run(6000, **_mc(sd, 90))
    weights = tr.surrogate_weights()
    tr.close()
    n_chains, samples, chunk, skip = 12, 200, 150, 40
    res = {"plain": [], "da": []}
    stats = None
    for mode in ("plain", "da"):
        for c in range(n_chains):
            eng = E.Engine(sd)
            if mode == "da":
                eng.surrogate_enable(screen=16, fill=0, train_every=0, seed=3)
                eng.surrogate_weights(weights)
            obs = []
            for i in range(samples):
                eng.mc_run(chunk, **_mc(sd, 7919 * (c + 1) + i))
                if i >= skip:
                    obs.append(_observables(eng, sd))
            res[mode].append(np.mean(obs, 0))
            if mode == "da":
                stats = eng.surrogate_stats()
                run, full = eng.totals(), eng.energy_full()
                for k in E.TERMS:
                    assert abs(run[k] - full[k]) <= 1e-9 * max(1.0, abs(full[k])), (k, run[k], full[k])
            eng.close()
    print("da stats (last chain)", stats, flush=True)
    p, d = np.array(res["plain"]), np.array(res["da"])
    err = lambda a: a.std(0, ddof=1) / np.sqrt(len(a))
    z = np.abs(p.mean(0) - d.mean(0)) / np.hypot(err(p), err(d))
    for j, label in enumerate(("energy", "Rg^2")):
        print("%s: plain %.4f +- %.4f, DA %.4f +- %.4f, |z| = %.2f" % (
            label, p.mean(0)[j], err(p)[j], d.mean(0)[j], err(d)[j], z[j]), flush=True)
    assert stats["s1_rejected"] > 0
    assert z.max() < 2.0


def test_training_reduces_the_loss_and_the_trained_network_gains():
    """Collect exact samples, train, and compare the delayed-acceptance chain
    with the untrained and the trained network: exact evaluations per
    accepted move (the cost the surrogate exists to cut)."""
    sd = systems.config("slab_small")
    eng = E.Engine(sd)
    eng.surrogate_enable(screen=16, fill=8, train_every=0, train_samples=64, ring=512, seed=8)
    st0 = eng.mc_run(1500, **_mc(sd, 5))
    s0 = eng.surrogate_stats()
    loss0 = eng.surrogate_train(1)
    loss1 = eng.surrogate_train(300)
    print("training loss %.3f -> %.3f" % (loss0, loss1), flush=True)
    assert loss1 < 0.5 * loss0
    st1 = eng.mc_run(1500, **_mc(sd, 6))
    s1 = eng.surrogate_stats()

    def per_accept(st, a, b):
        acc = sum(st.accepted)
        exact = (b["s1_passed"] - a["s1_passed"])
        return exact / max(acc, 1), acc

    base = {k: 0 for k in s0}
    e0, a0 = per_accept(st0, base, s0)
    e1, a1 = per_accept(st1, s0, s1)
    print("exact evaluations per accepted move: untrained %.2f (%d accepted), trained %.2f "
          "(%d accepted)" % (e0, a0, e1, a1), flush=True)
    assert a1 > 0 and e1 < e0
    eng.close()


def test_delayed_acceptance_over_a_batch_of_systems():
    """Four systems advanced together, all screened by the first system's
    network (one descriptor and one inference launch per batch for all of
    them), only the first feeding its training ring: every system keeps its
    running totals equal to its from-scratch energy."""
    sds = [systems.config("small", seed=50 + r) for r in range(4)]
    engs = [E.Engine(sd) for sd in sds]
    engs[0].surrogate_enable(screen=16, fill=2, train_every=8, train_samples=32, seed=3)
    b = E.Batch(engs)
    b.set_tuning(pmax=16)
    mc = [_mc(sd, 300 + r) for r, sd in enumerate(sds)]
    before = [e.positions() for e in engs]
    st = b.mc_run(2000, mc)
    s = engs[0].surrogate_stats()
    print("batched DA: %d moves, accepted %s, stats %s" % (st.moves, list(st.accepted), s),
          flush=True)
    assert st.moves == 4 * 2000
    assert s["s1_rejected"] > 0 and s["train_steps"] > 0
    for e, x0 in zip(engs, before):
        run, full = e.totals(), e.energy_full()
        for k in E.TERMS:
            if k == "dipole":
                continue
            assert abs(run[k] - full[k]) <= 1e-9 * max(1.0, abs(full[k])), (k, run[k], full[k])
        assert np.abs(e.positions() - x0).max() > 0
    b.close()
    for e in engs:
        e.close()


def test_brush_adapted_then_frozen_surrogate_saves_exact_work():
    """The schedule the bench runs on the confined brush (6,800 particles,
    slab Ewald, walls): the network adapts online from scratch, then is
    frozen.  From the same configuration the frozen two-stage chain and the
    plain chain each make 40,000 moves: the surrogate spends fewer than half
    the exact evaluations per accepted move that the plain chain does (every
    plain proposal is evaluated exactly), and its running totals stay equal
    to the from-scratch energy.  (The brush built by the workload generator
    is still relaxing after 3,000,000 plain moves, its energy falling
    steadily, so an ensemble comparison at test length would compare
    relaxation speeds; the ensemble check of the same schedule runs on the
    small slab.)"""
    sd = systems.config("c3")
    tr = E.Engine(sd)
    tr.surrogate_enable(screen=32, fill=4, train_every=16, train_samples=128, ring=2048, lr=1e-3, r_d=4.0,
                        seed=1)
    tr.set_tuning(pmax=32, cell_div=3)
    tr.mc_run(40000, **_mc(sd, 62))
    weights = tr.surrogate_weights()
    tr.close()
    per_acc = {}
    for mode in ("plain", "da"):
        eng = E.Engine(sd)
        if mode == "da":
            eng.surrogate_enable(screen=32, fill=0, train_every=0, train_samples=1, ring=1, r_d=4.0, seed=1)
            eng.surrogate_weights(weights)
            eng.set_tuning(pmax=32, cell_div=3)
        else:
            eng.set_tuning(pmax=16, cell_div=3)
        st = eng.mc_run(40000, **_mc(sd, 63))
        acc = sum(st.accepted)
        if mode == "da":
            s = eng.surrogate_stats()
            exact = s["s1_passed"] + s["train_evals"]
            run, full = eng.totals(), eng.energy_full()
            for k in E.TERMS:
                if k != "dipole":
                    assert abs(run[k] - full[k]) <= 1e-9 * max(1.0, abs(full[k])), (k, run[k], full[k])
        else:
            exact = sum(st.attempted)
        per_acc[mode] = exact / max(acc, 1)
        print("brush %s: %d accepted of %d moves, %.2f exact evaluations per accepted move" % (
            mode, acc, st.moves, per_acc[mode]), flush=True)
        eng.close()
    assert per_acc["da"] < 0.5 * per_acc["plain"]


def test_slab_adapted_then_frozen_surrogate_samples_the_plain_ensemble():
    """The same schedule (online adaptation, then a frozen network) on the
    slab with walls and the slab dipole term: energy and Rg^2 of twelve
    independent frozen delayed-acceptance chains agree with twelve plain
    chains within 2 sigma, all from one configuration equilibrated by a long
    plain run.  The dipole dE is the exact one here (dipole_exact = 1): with
    the reference's lag (potential_ewald.cc:528-531) a move's own dipole
    change is charged to the next proposal evaluated exactly, so a chain
    that evaluates fewer proposals exactly (the first stage rejects most)
    carries a different sequence of lagged terms and is not the same chain
    in distribution."""
    sd0 = systems.config("slab_small")
    eq = E.Engine(sd0, dipole_exact=1)
    eq.mc_run(100000, **_mc(sd0, 71))
    sd = dict(sd0, xyz=eq.positions())
    eq.close()
    tr = E.Engine(sd, dipole_exact=1)
    tr.surrogate_enable(screen=16, fill=2, train_every=8, train_samples=32, seed=4)
    tr.mc_run(6000, **_mc(sd, 72))
    weights = tr.surrogate_weights()
    tr.close()
    n_chains, samples, chunk, skip = 12, 200, 150, 40
    res = {"plain": [], "da": []}
    for mode in ("plain", "da"):
        for c in range(n_chains):
            eng = E.Engine(sd, dipole_exact=1)
            if mode == "da":
                eng.surrogate_enable(screen=16, fill=0, train_every=0, seed=4)
                eng.surrogate_weights(weights)
            obs = []
            for i in range(samples):
                eng.mc_run(chunk, **_mc(sd, 6151 * (c + 1) + i))
                if i >= skip:
                    obs.append(_observables(eng, sd))
            res[mode].append(np.mean(obs, 0))
            eng.close()
    p, d = np.array(res["plain"]), np.array(res["da"])
    err = lambda a: a.std(0, ddof=1) / np.sqrt(len(a))
    z = np.abs(p.mean(0) - d.mean(0)) / np.hypot(err(p), err(d))
    for j, label in enumerate(("energy", "Rg^2")):
        print("slab %s: plain %.4f +- %.4f, DA %.4f +- %.4f, |z| = %.2f" % (
            label, p.mean(0)[j], err(p)[j], d.mean(0)[j], err(d)[j], z[j]), flush=True)
    assert z.max() < 2.0
