"""CPU checks of the batched-replica workload builder: the native generator's
invariants and the builder's Ewald cutoff choice under the reference's own
cutoff rule (computed by the oracle)."""
import numpy as np
import pytest

from oracle import oracle as O
from plum_amd import host, systems


@pytest.mark.parametrize("rc", [3, 5, 8, 12])
def test_alpha_for_cutoff_lands_on_the_reference_rule(rc):
    a = systems.alpha_for_cutoff(rc)
    info = O.derive_cutoffs([40.0, 40.0, 40.0], 1.0, a, 0)
    assert info.real_cutoff == rc


def test_replica_specs_follow_the_workload_ranges():
    specs = [systems.replica_spec(r) for r in range(256)]
    assert specs == [systems.replica_spec(r) for r in range(256)]   # seeded
    assert all(10 <= s["chain_len"] <= 200 for s in specs)
    assert all(5 <= s["n_chains"] <= 100 for s in specs)
    assert all(0 <= s["salt_pairs"] <= 2000 for s in specs)
    assert all(0.02 <= s["rho"] <= 0.2 for s in specs)
    frac_charged = np.mean([s["charged"] for s in specs])
    assert 0.35 < frac_charged < 0.65
    total = sum(s["n"] for s in specs)
    assert 2.0e6 < total < 3.2e6      # ~10k particles per replica on average
    for s in specs:
        assert abs(s["n"] / s["box"] ** 3 - s["rho"]) < 1e-9 * s["rho"]


def _min_image_dist(a, b, L):
    d = a - b
    d -= L * np.round(d / L)
    return np.sqrt((d * d).sum(-1))


def test_native_generator_invariants():
    L = 14.0
    xyz, q, ty, ms = host.synth_system(6, 15, 90, 20, [L, L, L], seed=42)
    n = len(q)
    assert n == 6 * 15 + 90 + 40 and len(ms) == 6 + 90 + 40 + 1
    assert ms[0] == 0 and ms[-1] == n
    assert np.all(np.diff(ms)[:6] == 15) and np.all(np.diff(ms)[6:] == 1)
    assert abs(q.sum()) < 1e-12                      # counterions + neutral salt
    assert np.all(ty[:90] == 0) and np.all(ty[90:] == 1)
    for c in range(6):                              # bonds of exactly r0
        seg = xyz[ms[c]:ms[c + 1]]
        bl = np.linalg.norm(np.diff(seg, axis=0), axis=1)
        assert np.allclose(bl, 1.0, atol=1e-12)
    w = xyz.copy()
    dmin = min(_min_image_dist(w[i], w[i + 1:], L).min() for i in range(n - 1))
    assert dmin >= 0.8 - 1e-12                      # overlap rule
    again = host.synth_system(6, 15, 90, 20, [L, L, L], seed=42)
    assert np.array_equal(again[0], xyz)             # seeded


def test_replica_system_reduced_copies():
    sds = systems.replica_set(5, seed=3, n_max=700)
    for sd in sds:
        n = len(sd["q"])
        assert n <= 700 + 200
        L = sd["box"][0]
        assert L > 2 * 2.5
        assert abs(np.sum(sd["q"])) < 1e-9
