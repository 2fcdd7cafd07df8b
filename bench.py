"""Throughput of the trial-move energy engine on one MI355X.

A "step" is one device-resident sweep of `--moves-per-step` trial moves
(proposal, per-term dE, Metropolis decision, commit, all on the GPU) on the
40,000-particle bulk system: 200 chains x 100 beads (q = -1) + 20,000
counterions, density 0.05, Ewald alpha = 0.08 (r_c = 12, |n| <= 18), LJ
eps = sigma = 1 cut at 2.5, springs k = 400, r0 = 1, move mix 10% COM, 40%
pivot, 30% crankshaft, 20% reptation.  Inputs are resident in HBM before the
timed region.  `value` is the FP64 path; beside it: the FP32-pair path and
its per-term dE error, the confined brush with plain vs surrogate-screened
(delayed acceptance) moves, and the 256-replica batch.

Prints one JSON line (see the module docstring of each field below).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from plum_amd import engine as E  # noqa: E402
from plum_amd import systems  # noqa: E402

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0            # dense MFMA
MLP_FLOP_PER_ROW = 2.0 * (64 * 128 + 2 * 128 * 128 + 128)   # 64-128-128-128-1 network
# algorithmic FLOP per pair-image evaluation (minimum image, r, erfc, LJ)
FLOP_PER_PAIR = 45.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--moves-per-step", type=int, default=2000)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--profile-moves", type=int, default=300)
    ap.add_argument("--no-brush", action="store_true")
    ap.add_argument("--no-replicas", action="store_true")
    ap.add_argument("--brush-moves", type=int, default=20000)
    ap.add_argument("--replica-sweeps", type=int, default=512)
    ap.add_argument("--no-gc", action="store_true")
    ap.add_argument("--gc-steps", type=int, default=20000)
    ap.add_argument("--brush-only", action="store_true",
                    help="run only the brush / delayed-acceptance leg")
    ap.add_argument("--gc-only", action="store_true",
                    help="run only the grand-canonical leg (for profiling)")
    ap.add_argument("--growth-only", action="store_true",
                    help="only whole-chain growths of the grand-canonical workload (for profiling)")
    ap.add_argument("--brush-batched-only", action="store_true",
                    help="only the batched delayed-acceptance brushes, network at its initial "
                         "weights (one launch graph in the process, for kernel traces)")
    ap.add_argument("--replicas-only", action="store_true",
                    help="only the 256 replicas at a fixed launch shape (one launch graph, for kernel traces)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only device copies of known size (scale of the HBM counters)")
    return ap.parse_args()


def mc_kwargs(sd):
    return dict(move_prob=sd["move_prob"], move_size=sd["move_size"],
                phantom=sd.get("phantom", 0), grafted=sd.get("grafted", 0), bond_len=sd["r0"])


def timed_run(eng, sd, steps, warmup, mps, seed):
    kw = mc_kwargs(sd)
    for w in range(warmup):
        eng.mc_run(mps, seed=seed, **kw)
    st_tot = {"att": np.zeros(5), "acc": np.zeros(5), "moves": 0}
    t0 = time.perf_counter()
    for s in range(steps):
        st = eng.mc_run(mps, seed=seed, **kw)   # ends with a stream synchronize
        st_tot["att"] += np.array(st.attempted[:])
        st_tot["acc"] += np.array(st.accepted[:])
        st_tot["moves"] += st.moves
    t1 = time.perf_counter()
    return st_tot, t1 - t0


def real_space_work(sd, info):
    """Candidate pair-images examined per moved-bead query: the expected
    volume of the visited cells, (2 rs + h)^3 per dimension product, times
    the bead density."""
    n = len(sd["q"])
    L = np.array(info.box_e)
    vol = float(np.prod(L))
    rho = n / vol
    rs = max(info.real_cutoff, 2.5)
    h = L / np.array(info.n_cells)
    return rho * float(np.prod(2 * rs + h))


def real_space_union(sd, info, pos, n_sample=24, seed=3):
    """Distinct candidate slots one trial move's real-space queries read: the
    beads in the union of the cells the moved beads' queries visit (cells
    whose box lies within the search radius of a moved bead), averaged over
    sampled chains and the move mix (COM, pivot and reptation move the whole
    chain, crankshaft about a third of it)."""
    L = np.array(info.box_e)
    nc = np.array(info.n_cells)
    h = L / nc
    rs = max(info.real_cutoff, 2.5)
    w = pos - L * np.floor(pos / L)
    c3 = np.minimum((w / h).astype(int), nc - 1)
    flat = (c3[:, 0] * nc[1] + c3[:, 1]) * nc[2] + c3[:, 2]
    counts = np.bincount(flat, minlength=int(np.prod(nc)))
    ms = sd["mol_start"]
    chains = [m for m in range(len(ms) - 1) if ms[m + 1] - ms[m] > 1]
    rng = np.random.default_rng(seed)
    span = [np.arange(-int(np.ceil(rs / h[k])) - 1, int(np.ceil(rs / h[k])) + 2) for k in range(3)]
    off = np.stack(np.meshgrid(*span, indexing="ij"), -1).reshape(-1, 3)

    def union(beads):
        cells = set()
        for b in beads:
            u = np.floor(w[b] / h).astype(int) + off
            lo = u * h
            dd = np.maximum(0.0, np.maximum(lo - w[b], w[b] - (lo + h)))
            keep = u[(dd * dd).sum(1) <= rs * rs]
            cc = keep % nc
            cells.update(((cc[:, 0] * nc[1] + cc[:, 1]) * nc[2] + cc[:, 2]).tolist())
        return int(counts[list(cells)].sum())

    mp = sd["move_prob"]
    whole, third = [], []
    for m in rng.choice(chains, size=min(n_sample, len(chains)), replace=False):
        a, b = ms[m], ms[m + 1]
        whole.append(union(range(a, b)))
        k = max((b - a) // 3, 1)
        s0 = a + (b - a - k) // 2
        third.append(union(range(s0, s0 + k)))
    return ((mp[1] + mp[2] + mp[4]) * np.mean(whole) + mp[3] * np.mean(third)) / max(sum(mp[1:]), 1e-12)


def fp32_spec_errors(sd, moves=300):
    """Per-term dE error of the FP32-pair speculative kernel (k_compute_spec<1>,
    the timed loop): its recorded decisions replayed through the FP64 host path
    on the same trial coordinates; error / max(|dE|, 1 kT) and in kT."""
    e32 = E.Engine(sd, precision=E.PLUM_FP32_PAIR)
    e32.mc_run(200, seed=5, **mc_kwargs(sd))
    start = e32.positions()
    e32.mc_record(moves + 16)
    e32.mc_run(moves, seed=6, **mc_kwargs(sd))
    recs, _ = e32.mc_records(moves + 16)
    e32.close()
    e64 = E.Engine(dict(sd, xyz=start), precision=E.PLUM_FP64)
    worst = {k: 0.0 for k in E.TERMS}
    worst_abs = {k: 0.0 for k in E.TERMS}
    for r in recs:
        if not r["evaluated"]:
            continue
        t64, d64 = e64.delta(r["mol"], r["idx"], r["trial"])
        if t64 < 1e8 and r["total"] < 1e8:
            for k in E.TERMS:
                err = abs(r["dE"][k] - d64[k])
                worst_abs[k] = max(worst_abs[k], err)
                worst[k] = max(worst[k], err / max(abs(d64[k]), 1.0))
        e64.finalize(r["accepted"])
    e64.close()
    return ({k: float("%.3g" % v) for k, v in worst.items()},
            {k: float("%.3g" % v) for k, v in worst_abs.items()}, len(recs))


def cpu_baseline(sd, target_s=15.0):
    """The reference algorithm (oracle, FP64, one core): pairwise real space
    over all images plus the pairwise full-grid k-space sum, for ONE moved
    bead at its new position against a bounded prefix of the partners, then
    scaled to a full trial move (x2 for the old position, x mean moved beads,
    x N / partners)."""
    from oracle import oracle as O
    orc = O.Oracle(sd, init_energy=False)
    n = len(sd["q"])
    ms = sd["mol_start"]
    chains = [m for m in range(len(ms) - 1) if ms[m + 1] - ms[m] > 1]
    b0 = ms[chains[0]]
    p = sd["xyz"][b0] + np.array([0.1, 0.0, 0.0])
    parts = 8
    t = 0.0
    while True:
        t0 = time.perf_counter()
        orc.bead_energy_sample(p, sd["q"][b0], sd["type"][b0], parts)
        t = time.perf_counter() - t0
        log("cpu baseline: %d partners %.2f s" % (parts, t))
        if t > target_s / 8 or parts >= n:
            break
        parts = min(n, parts * 4)
    if t < target_s and parts < n:
        parts = min(n, int(parts * target_s / max(t, 1e-6)))
        t0 = time.perf_counter()
        orc.bead_energy_sample(p, sd["q"][b0], sd["type"][b0], parts)
        t = time.perf_counter() - t0
    # mean moved beads of the move mix: COM/pivot/reptation move all beads,
    # crankshaft about a third
    clen = ms[chains[0] + 1] - ms[chains[0]]
    mp = sd["move_prob"]
    n_moved = mp[0] * 1 + (mp[1] + mp[2] + mp[4]) * clen + mp[3] * clen / 3.0
    sec_per_move = t * (n / parts) * 2 * n_moved
    orc.close()
    return {
        "value": 1.0 / sec_per_move, "unit": "moves/s", "cores": 1, "kind": "port",
        "sample": ("oracle (reference algorithm, FP64, 1 thread): one moved bead vs %d of %d "
                   "partners took %.2f s; scaled x%.0f partners, x2 old/new, x%.1f mean moved "
                   "beads per move" % (parts, n, t, n / parts, n_moved)),
        "sample_seconds": t,
    }


def _cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_fast_baseline(sd, target_s=15.0):
    """The engine's own algorithm on the host cores (oracle/cpu_fast.c: FP64
    cell-list LJ + real space and a running S(k), OpenMP on every thread the
    process has): the trial moves a device run decided are evaluated and the
    accepted ones committed, so CPU and GPU do the same moves.  A calibration
    replay of a few moves sizes the timed replay to about target_s."""
    from oracle import oracle as O
    eng = E.Engine(sd, precision=E.PLUM_FP64)
    start = eng.positions().copy()
    cap = 6000
    eng.mc_record(cap)
    eng.mc_run(cap - 64, seed=3, **mc_kwargs(sd))
    recs, _ = eng.mc_records(cap)
    maxm = max(len(r["idx"]) for r in recs)
    eng.close()
    cal = O.CpuFast(dict(sd, xyz=start))
    t_cal, _ = cal.replay(recs[:40], maxm)
    cal.close()
    n = int(min(len(recs), max(40, 40 * target_s / max(t_cal, 1e-6))))
    cf = O.CpuFast(dict(sd, xyz=start))
    t, _ = cf.replay(recs[:n], maxm)
    cf.close()
    threads = O.CpuFast.threads()
    log("cpu_fast: %d moves in %.2f s on %d threads" % (n, t, threads))
    return {
        "value": n / t, "unit": "moves/s", "cores": threads, "kind": "port",
        "sample": ("cpu_fast (the engine's algorithm in C, FP64, OpenMP): the first %d of the "
                   "trial moves a device run of the same workload decided, evaluated (LJ + real "
                   "space by cell list, k space from a running S(k)) and the accepted ones "
                   "committed, in %.2f s" % (n, t)),
        "sample_seconds": t,
        "cpu": _cpu_model(), "nproc": os.cpu_count(),
        "flags": "gcc -O3 -fopenmp -std=gnu11 (generic x86-64 code)",
    }


def fp32_term_errors(sd, moves=20):
    """Per-term dE of the FP32-pair path against the FP64 path on identical
    chain moves (host path: the caller's trial coordinates)."""
    e64 = E.Engine(sd, precision=E.PLUM_FP64)
    e32 = E.Engine(sd, precision=E.PLUM_FP32_PAIR)
    full = e64.energy_full()
    rng = np.random.default_rng(7)
    ms = sd["mol_start"]
    chains = [m for m in range(len(ms) - 1) if ms[m + 1] - ms[m] > 1]
    worst = {k: 0.0 for k in E.TERMS}
    for _ in range(moves):
        mol = chains[rng.integers(len(chains))]
        a, b = ms[mol], ms[mol + 1]
        idx = np.arange(a, b, dtype=np.int32)
        tr = e64.positions()[a:b] + rng.normal(0, 0.05, size=(b - a, 3))
        t64, d64 = e64.delta(mol, idx, tr)
        _, d32 = e32.delta(mol, idx, tr)
        for k in E.TERMS:
            den = max(abs(d64[k]), 1e-3 * abs(full[k]), 1e-2)
            worst[k] = max(worst[k], abs(d32[k] - d64[k]) / den)
        acc = t64 < 1e8 and rng.random() < np.exp(-t64)
        e64.finalize(acc)
        e32.finalize(acc)
    e64.close()
    e32.close()
    return {k: float("%.3g" % v) for k, v in worst.items()}


def integrated_autocorr_time(x):
    """Integrated autocorrelation time tau = 1 + 2 sum_lag rho(lag) of a series
    (in samples), summed over lags until the window reaches 5 tau (Sokal's
    self-consistent window), after removing the series' linear trend (the
    brush is still relaxing at bench length: its energy falls steadily over
    millions of moves, and the trend alone would read as a long correlation)."""
    x = np.asarray(x, dtype=float)
    n = len(x)
    if n >= 3:
        t = np.arange(n, dtype=float)
        x = x - np.polyval(np.polyfit(t, x, 1), t)
    else:
        x = x - np.mean(x)
    var = np.dot(x, x) / n
    if n < 8 or var == 0:
        return 1.0
    tau = 1.0
    for lag in range(1, n // 2):
        tau += 2.0 * np.dot(x[:-lag], x[lag:]) / ((n - lag) * var)
        if lag >= 5 * tau:
            break
    return max(tau, 1.0)


def chain_ess(eng, moves, chunks, kw, seed):
    """Runs `moves` moves in `chunks` pieces, recording the total energy after
    each; returns (seconds, accepted, moves, ESS = chunks / tau of the energy
    series, tau)."""
    series, secs, acc, done = [], 0.0, 0, 0
    for c in range(chunks):
        st = eng.mc_run(moves // chunks, seed=seed + c, **kw)
        secs += st.seconds
        acc += sum(st.accepted)
        done += st.moves
        series.append(sum(eng.totals().values()))
    tau = integrated_autocorr_time(series)
    return secs, acc, done, len(series) / tau, tau


def brush_leg(moves, batch=32):
    """Confined brush (64 x 50 grafted chains + counterions, 400 charged
    surface sites, slab Ewald): the plain speculative path against delayed
    acceptance with the surrogate, in accepted moves per second and effective
    samples of the energy per second.  The surrogate adapts online from
    scratch (exact dE of proposals evaluated anyway as training data), then
    is frozen: diminishing adaptation taken to its limit, after which the
    first stage is a fixed function and the two-stage chain is exactly the
    Metropolis chain's target.  Frozen, it screens 32 proposals per batch and
    spends no exact slots on training.  Both arms take the exact slab dipole
    dE (dipole_exact): with the reference's lag a move's dipole change is
    charged to the next exactly evaluated proposal, and the surrogate chain
    evaluates far fewer proposals exactly, so the two would not sample the
    same chain."""
    sd = systems.config("c3")
    kw = dict(move_prob=sd["move_prob"], move_size=sd["move_size"], bond_len=sd["r0"],
              phantom=sd["phantom"], grafted=sd["grafted"])
    res = {"workload": "brush 64x50 grafted at 0.1/sigma^2 + 3200 ions + 400 surface sites, "
                       "walls z=0,40, slab Ewald (Lz 190), moves 20/50/30 COM/pivot/crankshaft",
           "particles": int(len(sd["q"])), "moves": moves}
    out = {}
    shapes = {"plain_p1": (1, 1), "plain_p8": (8, 1), "plain_p16": (16, 1), "plain_p8_cells3": (8, 3),
              "plain_p16_cells3": (16, 3)}
    for label, (pmax, cdiv) in shapes.items():
        eng = E.Engine(sd, dipole_exact=1)
        eng.set_tuning(pmax=pmax, cell_div=cdiv)
        eng.mc_run(2000, seed=1, **kw)
        st = eng.mc_run(moves, seed=2, **kw)
        acc = sum(st.accepted)
        out[label] = (st.moves / st.seconds, acc / st.seconds, acc / max(st.moves, 1))
        log("brush %s: %.0f moves/s, %.0f accepted/s" % (label, out[label][0], out[label][1]))
        eng.close()
    # the plain path at its best batch size and cell grid is the reference point
    best = max((k for k in shapes if k != "plain_p1"), key=lambda k: out[k][1])
    out["plain"] = out[best]
    plain_pmax, plain_div = shapes[best]
    # effective samples of the energy per second, plain
    eng = E.Engine(sd, dipole_exact=1)
    eng.set_tuning(pmax=plain_pmax, cell_div=plain_div)
    eng.mc_run(2000, seed=1, **kw)
    p_secs, p_acc, p_moves, p_ess, p_tau = chain_ess(eng, 4 * moves, 100, kw, 7)
    eng.close()
    # adaptation: online training from scratch
    eng = E.Engine(sd, dipole_exact=1)
    eng.surrogate_enable(screen=32, fill=4, train_every=16, train_samples=128, ring=2048, lr=1e-3,
                         r_d=4.0, seed=1)
    eng.set_tuning(pmax=32, cell_div=3)
    eng.mc_run(3 * moves, seed=1, **kw)
    s0 = eng.surrogate_stats()
    st = eng.mc_run(moves, seed=2, **kw)
    s1 = eng.surrogate_stats()
    acc = sum(st.accepted)
    adapt = (st.moves / st.seconds, acc / st.seconds, acc / max(st.moves, 1))
    log("brush DA adapting: %.0f moves/s, %.0f accepted/s" % (adapt[0], adapt[1]))
    # frozen: the trained weights, no training, no fill slots
    w = eng.surrogate_weights()
    loss, steps = s1["last_loss"], s1["train_steps"]
    eng.surrogate_enable(screen=32, fill=0, train_every=0, train_samples=1, ring=1, lr=1e-3, r_d=4.0, seed=1)
    eng.surrogate_weights(w)
    eng.set_tuning(pmax=32, cell_div=3)
    f0 = eng.surrogate_stats()
    f_secs, f_acc, f_moves, f_ess, f_tau = chain_ess(eng, 4 * moves, 100, kw, 11)
    f1 = eng.surrogate_stats()
    da = (f_moves / f_secs, f_acc / f_secs, f_acc / max(f_moves, 1))
    log("brush DA frozen: %.0f moves/s, %.0f accepted/s, ESS/s %.1f (plain %.1f)" % (
        da[0], da[1], f_ess / f_secs, p_ess / p_secs))
    # the surrogate kernels' rates: device durations (HIP events) of one batch's
    # descriptor and inference launches, and the rows a launch carries (2 rows
    # per moved bead of each screened proposal; the mean moved count from a
    # record of the decided proposals)
    eng.mc_record(4000)
    eng.set_profiling(True)
    eng.mc_run(3000, seed=13, **kw)
    kt = eng.kernel_times()
    eng.set_profiling(False)
    recs, _ = eng.mc_records(4000)
    eng.mc_record(0)
    n_bar = float(np.mean([r["n"] for r in recs])) if recs else 0.0
    rows = 2.0 * n_bar * 32
    d_us, m_us = kt["k_descr_spec"][0], kt["k_mlp_spec"][0]
    mlp_tf = rows * MLP_FLOP_PER_ROW / (m_us * 1e-6) / 1e12 if m_us > 0 else 0.0
    sur_k = {"k_descr_spec_us": d_us, "k_mlp_spec_us": m_us, "rows_per_launch": rows,
             "mean_moved_beads": n_bar, "descriptor_rows_per_s": rows / (d_us * 1e-6) if d_us > 0 else 0.0,
             "mlp": {"bound": "mfma", "achieved": mlp_tf, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": mlp_tf / PEAK_BF16_TFLOPS,
                     "basis": "%.0f FLOP per row (64-128-128-128-1) x %.0f rows per launch" % (
                         MLP_FLOP_PER_ROW, rows)}}
    log("brush surrogate kernels: descriptors %.1f us, inference %.1f us for %.0f rows (%.2f TFLOP/s)" % (
        d_us, m_us, rows, mlp_tf))
    # the batched form: several brushes advanced together, every one screened
    # by the frozen network (it stays with the first system)
    R = batch
    sds = [sd] + [systems.config("c3", seed=1003 + 101 * r) for r in range(1, R)]
    engs = [eng] + [E.Engine(x, dipole_exact=1) for x in sds[1:]]
    mcs = [dict(kw, seed=500 + r) for r in range(R)]
    # cells of r_s/3 and r_s/4: finer cells cut the descriptor scans, which
    # bound the batched form, and slow the exact queries; the faster is kept
    b = E.Batch(engs)
    da_by_div = {}
    for cdiv in (3, 4):
        b.set_tuning(pmax=32, batches_per_graph=16, cell_div=cdiv)
        b.mc_run(moves // R, mcs)
        st_d = b.mc_run(moves // R * 2, mcs)
        da_by_div[cdiv] = st_d
        log("brush DA batched x%d, cells r_s/%d: %.0f moves/s, %.0f accepted/s" % (
            R, cdiv, st_d.moves / st_d.seconds, sum(st_d.accepted) / st_d.seconds))
    da_div = max(da_by_div, key=lambda k: sum(da_by_div[k].accepted) / da_by_div[k].seconds)
    stb = da_by_div[da_div]
    accb = sum(stb.accepted)
    b.close()
    eng.surrogate_disable()
    b = E.Batch(engs)
    b.set_tuning(pmax=plain_pmax, cell_div=plain_div)
    b.mc_run(moves // R, mcs)
    stp = b.mc_run(moves // R * 2, mcs)
    accp = sum(stp.accepted)
    log("brush plain batched x%d: %.0f moves/s, %.0f accepted/s" % (R, stp.moves / stp.seconds,
                                                                      accp / stp.seconds))
    b.close()
    for e in engs:
        e.close()
    res.update({
        "systems_in_batch": R,
        "surrogate_batched_moves_per_s": stb.moves / stb.seconds,
        "surrogate_batched_accepted_per_s": accb / stb.seconds,
        "surrogate_batched_cell_div": da_div,
        "surrogate_batched_accepted_per_s_by_cell_div": {k: sum(v.accepted) / v.seconds
                                                         for k, v in da_by_div.items()},
        "plain_batched_accepted_per_s": accp / stp.seconds,
        "surrogate_batched_vs_plain": (accb / stb.seconds) / out["plain"][1],
        "surrogate_batched_vs_plain_batched": (accb / stb.seconds) / (accp / stp.seconds),
        "plain_fp64_moves_per_s": out["plain"][0], "plain_fp64_accepted_per_s": out["plain"][1],
        "plain_fp64_accept_ratio": out["plain"][2], "plain_fp64_proposals_per_batch": plain_pmax,
        "plain_fp64_cell_div": plain_div,
        "plain_accepted_per_s_by_shape": {k: out[k][1] for k in shapes},
        "plain_one_proposal_moves_per_s": out["plain_p1"][0],
        "plain_one_proposal_accepted_per_s": out["plain_p1"][1],
        "surrogate_moves_per_s": da[0], "surrogate_accepted_per_s": da[1],
        "surrogate_accept_ratio": da[2],
        "accepted_per_s_vs_plain": da[1] / out["plain"][1],
        "accepted_per_s_vs_one_proposal_plain": da[1] / out["plain_p1"][1],
        "ess_per_s_energy_plain": p_ess / p_secs, "ess_per_s_energy_surrogate": f_ess / f_secs,
        "ess_per_s_vs_plain": (f_ess / f_secs) / (p_ess / p_secs),
        "autocorr_time_energy_chunks": {"plain": p_tau, "surrogate": f_tau,
                                        "chunk_moves": 4 * moves // 100},
        "first_stage_pass_fraction": (f1["s1_passed"] - f0["s1_passed"]) / max(f_moves, 1),
        "exact_evals_per_accepted": ((f1["s1_passed"] - f0["s1_passed"]) +
                                     (f1["train_evals"] - f0["train_evals"])) / max(f_acc, 1),
        "stale_cuts": f1["stale_cuts"] - f0["stale_cuts"],
        "surrogate_kernels": sur_k,
        "adapting_moves_per_s": adapt[0], "adapting_accepted_per_s": adapt[1],
        "adapting_exact_evals_per_accepted": ((s1["s1_passed"] - s0["s1_passed"]) +
                                              (s1["train_evals"] - s0["train_evals"])) / max(acc, 1),
        "training_loss_kT2": loss, "training_steps": steps,
        "schedule": "adapt online for 3 x moves (fill 4, a training step every 16 batches), then "
                    "freeze (no training, no fill slots, 32 proposals screened per batch); cells of "
                    "a third of the search radius for the surrogate runs",
    })
    return res


def brush_batched_leg(moves, R=32):
    """32 brushes advanced together with delayed acceptance, screened by the
    first system's network at its initial weights (no training: one launch
    graph in the process, which a kernel trace needs); the descriptor and
    inference kernels' work does not depend on the weights."""
    sd = systems.config("c3")
    kw = dict(move_prob=sd["move_prob"], move_size=sd["move_size"], bond_len=sd["r0"],
              phantom=sd["phantom"], grafted=sd["grafted"])
    sds = [sd] + [systems.config("c3", seed=1003 + 101 * r) for r in range(1, R)]
    engs = [E.Engine(x, dipole_exact=1) for x in sds]
    engs[0].surrogate_enable(screen=32, fill=0, train_every=0, train_samples=1, ring=1, lr=1e-3, r_d=4.0,
                             seed=1)
    mcs = [dict(kw, seed=500 + r) for r in range(R)]
    b = E.Batch(engs)
    b.set_tuning(pmax=32, batches_per_graph=16, cell_div=3)
    b.mc_run(max(1, moves // R // 4), mcs)
    st = b.mc_run(max(1, moves // R), mcs)
    s = engs[0].surrogate_stats()
    res = {"systems": R, "moves": st.moves, "moves_per_s": st.moves / st.seconds,
           "accepted_per_s": sum(st.accepted) / st.seconds,
           "first_stage_pass_fraction_system0": s["s1_passed"] / max(s["s1_passed"] + s["s1_rejected"], 1)}
    log("brush batched x%d (initial weights): %.0f moves/s" % (R, res["moves_per_s"]))
    b.close()
    for e in engs:
        e.close()
    return res


def move_work(sd, info):
    """Algorithmic FLOP of one trial move: the k-space update 16 n N_k + 8 N_k
    and the real-space pair sum 45 FLOP per candidate pair-image for the old
    and new position of each of the n moved beads (n: the move mix's mean;
    COM, pivot and reptation move the whole chain, crankshaft a third)."""
    ms = sd["mol_start"]
    clen = int(max(np.diff(ms)))
    mp = sd["move_prob"]
    n_moved = mp[0] * 1 + (mp[1] + mp[2] + mp[4]) * clen + mp[3] * max(clen - 2, 0) / 3.0
    kflop = 16.0 * n_moved * info.n_k_half + 8.0 * info.n_k_half
    return kflop + FLOP_PER_PAIR * 2 * n_moved * real_space_work(sd, info)


def replica_leg(sweeps, tuned=True):
    """256 independent bulk systems (chain length 10..200, 5..100 chains,
    0..2000 salt pairs, half of them charged, density 0.02..0.2), FP32 pair
    math, advanced together by the batched kernels after the launch-shape
    autotuner (tuned=False: the shape the autotuner picks here, fixed, so a
    kernel trace sees one launch graph).  The work rate is the algorithmic
    FLOP of the moves made over the wall time of the run, a lower bound on the
    kernels' rate."""
    t0 = time.perf_counter()
    sds = systems.replica_set(256)
    engs = [E.Engine(sd, precision=E.PLUM_FP32_PAIR) for sd in sds]
    log("replicas: %d particles ready in %.1f s" % (sum(len(s["q"]) for s in sds),
                                                    time.perf_counter() - t0))
    work = [move_work(sd, e.info()) for sd, e in zip(sds, engs)]
    b = E.Batch(engs)
    mc = [dict(move_prob=sd["move_prob"], move_size=1.0, seed=1000 + r, bond_len=1.0)
          for r, sd in enumerate(sds)]
    if tuned:
        t = b.tune(mc, moves_per_trial=16)
    else:
        b.set_tuning(blocks_per_cu=16, split_div=768, pmax=4)
        t = b.tuning()
        b.mc_run(max(1, sweeps // 8), mc)
    st, per = b.mc_run(sweeps, mc, per_replica=True)
    flop = sum(p.moves * w for p, w in zip(per, work))
    rate = flop / st.seconds / 1e12
    log("replicas: %d moves in %.2f s, %.2f TFLOP/s of algorithmic work (wall clock)" % (
        st.moves, st.seconds, rate))
    res = {"workload": "256 bulk replicas, chain length U[10,200] x U[5,100] chains, salt "
                       "U[0,2000] pairs, charged with p=0.5, density U[0.02,0.2]; FP32 pairs",
           "particles": int(sum(len(s["q"]) for s in sds)), "moves_per_replica": sweeps,
           "moves_per_s": st.moves / st.seconds,
           "accept_ratio": sum(st.accepted) / max(sum(st.attempted), 1),
           "tuning": {"blocks_per_cu": t.blocks_per_cu, "split_div": t.split_div, "pmax": t.pmax},
           "work_rate": {"achieved": rate, "unit": "TFLOP/s", "peak_fp64": PEAK_FP64_TFLOPS,
                         "frac_of_fp64_peak": rate / PEAK_FP64_TFLOPS,
                         "basis": "sum over replicas of moves x (16 n N_k + 8 N_k + 45 x 2 n candidates) "
                                  "FLOP over the run's wall time (kernel time is shorter)"}}
    b.close()
    for e in engs:
        e.close()
    return res


def growth_leg(n=40):
    """Whole 32-bead insertion growths (16 trials per bead) into the
    grand-canonical workload's starting configuration, with fixed draws:
    the device time of one growth."""
    sd = systems.config("c4")
    eng = E.Engine(sd)
    rng = np.random.default_rng(1)
    L, K = sd["gc"]["chain_len"], sd["gc"]["trials"]
    box = np.asarray(sd["box"])
    first = rng.uniform(0.1, 0.9, (2, 3)) * box
    v = rng.normal(size=(L, K, 3))
    disp = 0.98 * v / np.linalg.norm(v, axis=2)[..., None]
    ions = rng.uniform(0, 1, (L, K, 3)) * box
    sel = rng.uniform(0, 1, L)
    kw = dict(first=first, disp=disp, ions=ions, sel=sel, q_ion=1.0)
    for _ in range(3):
        eng.cbmc_grow(0, L, K, 0, -1.0, **kw)
    t = time.perf_counter()
    for _ in range(n):
        w, stop, _ = eng.cbmc_grow(0, L, K, 0, -1.0, **kw)
    dt = (time.perf_counter() - t) / n
    eng.close()
    log("growth: %.1f us per 32-bead growth (%.1f us per step)" % (1e6 * dt, 1e6 * dt / L))
    return {"growth_us": 1e6 * dt, "step_us": 1e6 * dt / L, "growths": n, "weight": w,
            "stop": stop}


def gc_leg(steps):
    """Grand-canonical bulk: 32-bead chains (charge -1 per bead) exchanged
    together with 32 counterions by configurational-bias insertion and
    deletion (16 trials per grown bead, each growth step's trial energies in
    one GPU pass), one exchange every 10 steps on average and translational
    moves on the device loop in between; starting from 50 chains in a box of
    60 at the chemical potential where 50 chains are in balance.  Reports the
    mean chain count with its block error, the drift between the halves, the
    exchange acceptance both ways and the volume-perturbation pressure
    (Yethiraj and de Miguel estimators, pressure.cc:187-387) sampled along
    the run."""
    import tempfile
    from plum_amd import host as H
    sd = systems.config("c4")
    g = sd["gc"]
    path = systems.write_reference_inputs(sd, tempfile.mkdtemp(prefix="plum_gc_"))
    sim = H.HostSimulation(path, 2024)
    # the reference's exchange, recording ln(W / w_1) of every insertion
    # growth for the two-stage run below (one extra first-bead trial energy
    # per growth; no random draw, the chain is the reference's)
    sim.gc_prescreen(1)
    per = 2 * g["chain_len"]
    t = time.perf_counter()
    w = [sim.widom() for _ in range(2)]
    widom_s = (time.perf_counter() - t) / 2
    growth_steps = 50 * (g["chain_len"] - 1)
    a0, c0 = sim.gc_stats()[0], sim.gc_stats()[1]
    att0, acc0, sec0 = sim.run_device(0)
    chains, py, pd = [], None, None
    t = time.perf_counter()
    nchunk = 50
    chunk = max(steps // nchunk, 1)
    done = 0
    t_p = 0.0
    while done < steps:
        att, acc, sec = sim.run_device(chunk)
        done += chunk
        chains.append(sim.gc_stats()[2] / per)
        tp = time.perf_counter()
        _, py, pd, n_p = sim.pressure_sample()
        t_p += time.perf_counter() - tp
    secs = time.perf_counter() - t - t_p
    ga, gacc, nb, nm = sim.gc_stats()
    ins = ga[0] - a0[0]
    dele = ga[1] - a0[1]
    moves = sum(att) - sum(att0)
    ch = np.array(chains)
    nb10 = len(ch) // 10 * 10
    blocks = ch[:nb10].reshape(10, -1).mean(1)
    half = len(ch) // 2
    log("gc: %d steps in %.2f s: %d insertions (%d accepted), %d deletions (%d accepted), "
        "<chains> %.2f +- %.2f, P %.3e / %.3e" % (done, secs, ins, gacc[0] - c0[0], dele, gacc[1] - c0[1],
                                                ch.mean(), blocks.std(ddof=1) / np.sqrt(10), py, pd))
    res = {"workload": "grand-canonical bulk, 32-bead chains (-1 per bead) + 32 counterions "
                       "exchanged by configurational bias, 16 trials per bead, mu %.1f kT, box 60, "
                       "start 50 chains, exchange every 10 steps" % g["chem_pot"],
           "steps": done, "steps_per_s": done / secs,
           "exchanges_per_s": (ins + dele) / secs,
           "insertions": ins, "insertions_accepted": gacc[0] - c0[0],
           "deletions": dele, "deletions_accepted": gacc[1] - c0[1],
           "translational_moves_per_s": moves / max(sec[0] - sec0[0], 1e-9),
           "seconds_in_exchanges": sec[1] - sec0[1], "seconds_in_device_loop": sec[0] - sec0[0],
           "chains_mean": float(ch.mean()), "chains_err": float(blocks.std(ddof=1) / np.sqrt(10)),
           "chains_std": float(ch.std(ddof=1)),
           "chains_first_half": float(ch[:half].mean()), "chains_second_half": float(ch[half:].mean()),
           "pressure_yethiraj": py, "pressure_demiguel": pd, "pressure_samples": n_p,
           "pressure_unit": "kT per unit volume (volume perturbation dz = 1e-5 along z)",
           "growth_step_us": 1e6 * widom_s / growth_steps,
           "widom_mean_weight": float(np.mean(w))}
    _, _, mean_log_rest, _, n_rest = sim.gc_prescreen_stats()
    sim.close()
    res["two_stage"] = gc_two_stage(path, steps, mean_log_rest, n_rest, per, sec[1] - sec0[1], ins + dele)
    res["two_stage_reference_deck"] = gc_two_stage_deck(os.path.join(ROOT, "tests", "golden", "examples",
                                                                      "bulk_muvt", "run.in"), 4000)
    return res


def gc_two_stage_deck(path, steps, seed=7):
    """The reference's own grand-canonical deck (8-bead chains of charge -1
    with counterions, 30 trials, mu = -90, box 100): the reference's exchange
    recording ln(W / w_1), then the two-stage exchange with ln theta their
    mean, the same number of steps from the same start; growths run and
    exchanges accepted by each."""
    from plum_amd import host as H
    out = {}
    log_theta = 0.0
    for mode in (1, 2):
        sim = H.HostSimulation(path, seed)
        sim.gc_prescreen(mode, log_theta)
        att, acc, sec = sim.run_device(steps)
        ga, gacc, nb, nm = sim.gc_stats()
        rej, pas, mean_log, _, n_rest = sim.gc_prescreen_stats()
        growths = sum(ga) if mode == 1 else sum(pas)
        out["reference" if mode == 1 else "two_stage"] = {
            "exchanges": ga, "accepted": gacc, "growths": growths, "first_stage_rejected": rej,
            "seconds_in_exchanges": sec[1], "beads_at_end": nb}
        if mode == 1:
            log_theta = mean_log
        sim.close()
    out["log_theta"] = log_theta
    r, t = out["reference"], out["two_stage"]
    log("gc reference deck: reference %d growths, accepted %s, %.2f s in exchanges; two-stage (log theta "
        "%.2f) %d growths, accepted %s, %.2f s" % (r["growths"], r["accepted"], r["seconds_in_exchanges"],
                                                  log_theta, t["growths"], t["accepted"],
                                                  t["seconds_in_exchanges"]))
    return out


def gc_two_stage(path, steps, log_theta, n_rest, per, plain_exchange_s, plain_exchanges):
    """The same run with the two-stage exchange (plum_sim_gc_prescreen mode
    2): first stage on the first bead's exact weight times theta, ln theta
    the mean of ln(W / w_1) over the growths the reference's run recorded,
    the growth and the exact Rosenbluth weight only past it."""
    from plum_amd import host as H
    sim = H.HostSimulation(path, 2024)
    sim.gc_prescreen(2, log_theta)
    att0, acc0, sec0 = sim.run_device(0)
    a0, c0 = sim.gc_stats()[0], sim.gc_stats()[1]
    chains = []
    nchunk = 50
    chunk = max(steps // nchunk, 1)
    done = 0
    while done < steps:
        att, acc, sec = sim.run_device(chunk)
        done += chunk
        chains.append(sim.gc_stats()[2] / per)
    ga, gacc, nb, nm = sim.gc_stats()
    rej, pas, _, _, _ = sim.gc_prescreen_stats()
    sim.close()
    ch = np.array(chains)
    nb10 = len(ch) // 10 * 10
    blocks = ch[:nb10].reshape(10, -1).mean(1)
    ex_s = sec[1] - sec0[1]
    exchanges = (ga[0] - a0[0]) + (ga[1] - a0[1])
    res = {"log_theta": log_theta, "growths_fitted": n_rest, "steps": done,
           "insertions": ga[0] - a0[0], "insertions_accepted": gacc[0] - c0[0],
           "deletions": ga[1] - a0[1], "deletions_accepted": gacc[1] - c0[1],
           "first_stage_rejected": rej, "first_stage_passed": pas,
           "seconds_in_exchanges": ex_s, "exchanges_per_s_in_exchanges": exchanges / max(ex_s, 1e-9),
           "plain_exchanges_per_s_in_exchanges": plain_exchanges / max(plain_exchange_s, 1e-9),
           "chains_mean": float(ch.mean()), "chains_err": float(blocks.std(ddof=1) / np.sqrt(10))}
    log("gc two-stage (log theta %.2f): %d exchanges in %.2f s (plain %d in %.2f s), first stage "
        "rejected %s passed %s, accepted %s, <chains> %.2f +- %.2f" % (
            log_theta, exchanges, ex_s, plain_exchanges, plain_exchange_s, rej, pas,
            [gacc[0] - c0[0], gacc[1] - c0[1]], res["chains_mean"], res["chains_err"]))
    return res


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_compute_spec_traffic.json")


def measured_traffic():
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE per dispatch, each
    scaled by the same counter on a copy of known size), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
        return t["bytes_per_launch"], t
    except (OSError, ValueError, KeyError):
        return None, None


def copy_probe():
    """Device-to-device copies of 256 MiB: 256 MiB read and 256 MiB written
    per copy kernel, the known traffic the HBM counters are scaled by."""
    import torch
    n = 256 << 20
    a = torch.ones(n // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_each_way": n, "copies": 5}), flush=True)


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build measures a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path).", file=sys.stderr)
        sys.exit(2)
    if a.copy_probe:
        copy_probe()
        return
    if a.gc_only:
        print(json.dumps({"grand_canonical": gc_leg(a.gc_steps)}), flush=True)
        return
    if a.brush_only:
        print(json.dumps({"brush_delayed_acceptance": brush_leg(a.brush_moves)}), flush=True)
        return
    if a.growth_only:
        print(json.dumps({"growth": growth_leg()}), flush=True)
        return
    if a.replicas_only:
        print(json.dumps({"replicas": replica_leg(a.replica_sweeps, tuned=False)}), flush=True)
        return
    if a.brush_batched_only:
        print(json.dumps({"brush_batched": brush_batched_leg(a.brush_moves)}), flush=True)
        return
    t0 = time.perf_counter()
    sd = systems.config(a.config)
    n = len(sd["q"])
    log("system %s: %d particles built in %.1f s" % (a.config, n, time.perf_counter() - t0))
    eng = E.Engine(sd, precision=E.PLUM_FP64)
    info = eng.info()
    log("engine ready (FP64): %d half-space k, cells %s cap %d" % (
        info.n_k_half, list(info.n_cells), info.cell_cap))
    stats, secs = timed_run(eng, sd, a.steps, a.warmup, a.moves_per_step, seed=1)
    log("FP64: %d moves in %.3f s" % (stats["moves"], secs))
    moves = stats["moves"]
    value = moves / secs
    acc_ratio = float(stats["acc"].sum() / max(stats["att"].sum(), 1))

    # per-kernel device durations (HIP events on the engine's stream)
    eng.set_profiling(True)
    stp = eng.mc_run(a.profile_moves, seed=2, **mc_kwargs(sd))
    kt = eng.kernel_times()
    eng.set_profiling(False)
    # mean moved beads of the move mix (for the algorithmic work per move)
    ms = sd["mol_start"]
    clen = int(max(np.diff(ms)))
    mp = sd["move_prob"]
    n_moved = mp[0] * 1 + (mp[1] + mp[2] + mp[4]) * clen + mp[3] * max(clen - 2, 0) / 3.0
    cand = real_space_work(sd, info)
    # algorithmic work of one trial move in k_compute_spec: the k-space update
    # 16 n N_k + 8 N_k FLOP plus the real-space pair sum 45 FLOP per candidate
    # pair-image, old and new position of every moved bead; a launch carries a
    # speculative batch of moves, so the rate is total work over total time
    kflop = 16.0 * n_moved * info.n_k_half + 8.0 * info.n_k_half
    rflop = FLOP_PER_PAIR * 2 * n_moved * cand
    rl_name = "k_compute_spec"
    avg_us, launches = kt[rl_name]
    per_move = (kflop + rflop) / 1e12
    achieved = stp.moves * per_move / (avg_us * launches * 1e-6) if avg_us > 0 else 0.0
    per_launch = stp.moves * per_move / max(launches, 1)
    # algorithmic HBM bytes per move: k space 40 N_k (S(k) read, dS write,
    # c_k read), real space 20 B per distinct candidate slot
    n_union = real_space_union(sd, info, eng.positions())
    alg_bytes = 40.0 * info.n_k_half + 20.0 * n_union

    traffic, tinfo = measured_traffic()
    out = {
        "metric": "MC trial moves/s on 1 MI355X (plain vs surrogate-screened), \u0394E err vs FP64",
        "value": value, "unit": "moves/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1000.0 * secs / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "bulk NVT 200x100 beads + 20000 ions, rho 0.05, alpha 0.08 "
                               "(r_c 12, %d half-space k), %d moves per step" % (
                                   info.n_k_half, a.moves_per_step),
                   "particles": n, "moves_per_step": a.moves_per_step},
        "accept_ratio": acc_ratio,
        "kernel_avg_us": {k: round(v[0], 3) for k, v in kt.items()},
        "roofline": {"bound": "valu", "pipe": "FP64 VALU (dense FP64 vector peak)", "kernel": rl_name,
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "work_per_launch": per_launch, "moves_per_launch": stp.moves / max(launches, 1),
                     "algorithmic_bytes_per_move": alg_bytes,
                     "algorithmic_bytes_basis": "40 B x %d k + 20 B x %.0f distinct candidate slots" % (
                         info.n_k_half, n_union),
                     "hbm_achieved_gbs": alg_bytes * stp.moves / (avg_us * launches * 1e-6) / 1e9
                     if avg_us > 0 else 0.0,
                     "basis": "per move: k-space 16 x %.1f moved beads x %d k + 8 x %d k; real "
                              "space %.0f FLOP x 2 positions x %.1f moved beads x %.0f candidate "
                              "pair-images; x moves per launch" % (
                                  n_moved, info.n_k_half, info.n_k_half, FLOP_PER_PAIR, n_moved,
                                  cand)},
    }
    if tinfo is not None:
        out["roofline"]["traffic_source"] = tinfo.get("source")
        out["roofline"]["traffic_per_move"] = tinfo.get("bytes_per_move")
    if not a.no_fp32:
        e32 = E.Engine(sd, precision=E.PLUM_FP32_PAIR)
        st32, s32 = timed_run(e32, sd, a.steps, a.warmup, a.moves_per_step, seed=1)
        e32.set_profiling(True)
        e32.mc_run(a.profile_moves, seed=2, **mc_kwargs(sd))
        kt32 = e32.kernel_times()
        out["fp32pair_moves_per_s"] = st32["moves"] / s32
        log("FP32-pair: %d moves in %.3f s" % (st32["moves"], s32))
        out["fp32pair_kernel_avg_us"] = {k: round(v[0], 3) for k, v in kt32.items()}
        e32.close()
    eng.close()
    rel, ab, nrec = fp32_spec_errors(sd)
    out["dE_err_fp32pair_spec_kernel"] = {
        "rel_to_max_dE_1kT": rel, "abs_kT": ab, "decisions": nrec,
        "basis": "k_compute_spec<1> recorded decisions replayed through the FP64 host path"}
    out["dE_rel_err_fp32pair_vs_fp64"] = fp32_term_errors(sd)
    if not a.no_brush:
        out["brush_delayed_acceptance"] = brush_leg(a.brush_moves)
    if not a.no_replicas:
        out["replicas"] = replica_leg(a.replica_sweeps)
    if not a.no_gc:
        out["grand_canonical"] = gc_leg(a.gc_steps)
    if not a.no_cpu_baseline:
        # the CPU baseline is the engine's algorithm on every host core; the
        # reference's own algorithm on one core is reported beside it
        out["cpu_baseline"] = cpu_fast_baseline(sd)
        out["cpu_baseline"]["reference_algorithm_1core"] = cpu_baseline(sd)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
