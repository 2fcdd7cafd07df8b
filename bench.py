"""Throughput of the trial-move energy engine on one MI355X.

A "step" is one device-resident sweep of `--moves-per-step` trial moves
(proposal, per-term dE, Metropolis decision, commit, all on the GPU) on the
40,000-particle bulk system: 200 chains x 100 beads (q = -1) + 20,000
counterions, density 0.05, Ewald alpha = 0.08 (r_c = 12, |n| <= 18), LJ
eps = sigma = 1 cut at 2.5, springs k = 400, r0 = 1, move mix 10% COM, 40%
pivot, 30% crankshaft, 20% reptation.  Inputs are resident in HBM before the
timed region.  `value` is the FP64 path; the FP32-pair path is reported
beside it.

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
# algorithmic FLOP per pair-image evaluation (minimum image, r, erfc, LJ)
FLOP_PER_PAIR = 45.0


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--moves-per-step", type=int, default=1000)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--profile-moves", type=int, default=300)
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


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: this build measures a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path).", file=sys.stderr)
        sys.exit(2)
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
        "roofline": {"bound": "mfma", "pipe": "FP64 VALU (dense FP64 peak)", "kernel": rl_name,
                     "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "work_per_launch": per_launch, "moves_per_launch": stp.moves / max(launches, 1),
                     "basis": "per move: k-space 16 x %.1f moved beads x %d k + 8 x %d k; real "
                              "space %.0f FLOP x 2 positions x %.1f moved beads x %.0f candidate "
                              "pair-images; x moves per launch" % (
                                  n_moved, info.n_k_half, info.n_k_half, FLOP_PER_PAIR, n_moved,
                                  cand)},
    }
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
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sd)
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
