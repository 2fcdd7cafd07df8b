"""CPU checks that pin the oracle's grand-canonical chain exchange
(orc_gc_insertion / orc_gc_deletion / orc_gc_widom, restating
cbmc.cc:213-533) by what statistical mechanics fixes: with every
interaction switched off each trial weight is 1, so the Rosenbluth weight is
1 and the exchange must sample the ideal grand-canonical distribution.  For
neutral chains that is Poisson with mean z V L^1.5 / Lambda (z = e^{beta mu},
L the chain length standing in for the mass); with L counterions per chain,
P(N) ~ a^N / (N! (L N)!) with a = z V L^1.5 / Lambda (V / Lambda)^L.  A
wrong factorial, volume or prefactor in either acceptance rule moves the
sampled mean away.  The test insertion of the chemical potential returns its
prefactor exactly when every weight is 1."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from plum_amd import systems


def _ideal(L, charged, mu, n_start=2, box=10.0, seed=3, debroglie=1.0):
    sd = systems.config("tiny")
    xyz, q, ty, ms = systems.build_bulk(n_start, L, n_start * L if charged else 0, box, seed,
                                        bead_q=-1.0 if charged else 0.0)
    if charged:
        xyz, q, ty, ms = systems.gc_layout(xyz, q, ty, ms, n_start, L)
        ty = np.zeros_like(ty)
    sd.update(box=[box] * 3, xyz=xyz, q=q, type=ty, mol_start=ms, use_pair=0, use_ewald=0,
              gc=dict(chain_len=L, bead_charge=-1.0 if charged else 0.0, trials=4, freq=1,
                      chem_pot=mu, debroglie=debroglie))
    return sd


def _sample(sd, steps, seed):
    g = O.GCSim(sd, seed)
    L = sd["gc"]["chain_len"]
    per = 1 + (L if sd["gc"]["bead_charge"] != 0 else 0)
    ns = []
    for _ in range(steps):
        g.run(1)
        ns.append((len(g.sd["mol_start"]) - 1) // per)
    acc = g.gc_accepted
    g.close()
    ns = np.array(ns[steps // 10:], dtype=float)
    blocks = ns[: len(ns) // 20 * 20].reshape(20, -1).mean(1)
    return ns.mean(), blocks.std(ddof=1) / math.sqrt(20), acc


def test_ideal_neutral_chains_are_poisson():
    L, V, lam = 4, 1000.0, 4.0
    mu = math.log(lam / (V * L ** 1.5))
    mean, err, acc = _sample(_ideal(L, False, mu), 40000, 9)
    print("neutral: <N> = %.3f +- %.3f (exact %.3f), accepted %s" % (mean, err, lam, acc))
    assert min(acc) > 1000
    assert abs(mean - lam) < 4 * err + 0.02


def test_ideal_chains_with_counterions():
    L, V = 2, 1000.0
    mu = -19.0
    a = math.exp(mu) * V * L ** 1.5 * V ** L
    n = np.arange(0, 60)
    logp = n * math.log(a) - np.array([math.lgamma(k + 1) + math.lgamma(L * k + 1) for k in n])
    p = np.exp(logp - logp.max())
    exact = (n * p).sum() / p.sum()
    mean, err, acc = _sample(_ideal(L, True, mu), 40000, 5)
    print("charged: <N> = %.3f +- %.3f (exact %.3f), accepted %s" % (mean, err, exact, acc))
    assert 1.0 < exact < 10.0 and min(acc) > 1000
    assert abs(mean - exact) < 4 * err + 0.02


@pytest.mark.parametrize("charged", [False, True])
def test_widom_prefactor_of_ideal_chains(charged):
    L, V = 3, 1000.0
    sd = _ideal(L, charged, 0.0, n_start=2, debroglie=0.5)
    g = O.GCSim(sd, 1)
    w = g.widom(50)
    g.close()
    n_chain, n_ion = 2, (2 * L if charged else 0)
    want = (V / 15.625) / (0.5 / L ** 1.5) / (n_chain + 1)
    if charged:
        want *= ((V / 15.625) / 0.5) ** L / math.prod(range(n_ion + 1, n_ion + L + 1))
    assert w == pytest.approx(want, rel=1e-12)


def test_gasdev_pairs_and_moments():
    r = O.MT19937(4)
    gs = O.OrcGauss()
    x = np.array([O.lib().orc_gasdev(1.0, 0.5, O.C.byref(r.s), O.C.byref(gs)) for _ in range(40000)])
    assert abs(x.mean() - 1.0) < 0.01 and abs(x.std() - 0.5) < 0.01
    # the polar method draws two uniforms per pair: the second deviate of a
    # pair consumes none
    r1, r2 = O.MT19937(7), O.MT19937(7)
    g1 = O.OrcGauss()
    O.lib().orc_gasdev(0.0, 1.0, O.C.byref(r1.s), O.C.byref(g1))
    state = bytes(r1.s)
    O.lib().orc_gasdev(0.0, 1.0, O.C.byref(r1.s), O.C.byref(g1))
    assert bytes(r1.s) == state


@pytest.mark.parametrize("log_theta", [-1.0, 1.5])
def test_two_stage_exchange_of_ideal_chains_is_poisson(log_theta):
    """The two-stage exchange (orc_gc_insertion2 / orc_gc_deletion2: a first
    stage on the first bead's weight times theta, the second on the whole
    Rosenbluth weight) with every interaction off: whatever theta, the ideal
    chains' count stays Poisson with the same mean as the reference's
    exchange (a first stage that is not the exact inverse between insertion
    and deletion, or a second stage that does not divide it out, shifts it),
    and theta moves the work between the stages: exchanges stopped at the
    first stage appear exactly when theta makes r_1 < 1."""
    L, V, lam = 4, 1000.0, 4.0
    mu = math.log(lam / (V * L ** 1.5))
    sd = _ideal(L, False, mu)
    g = O.GCSim(sd, 19, log_theta=log_theta)
    ns = []
    steps = 40000
    for _ in range(steps):
        g.run(1)
        ns.append(len(g.sd["mol_start"]) - 1)
    rej, pas, acc = g.s1_rejected, g.s1_passed, g.gc_accepted
    g.close()
    ns = np.array(ns[steps // 10:], dtype=float)
    blocks = ns[: len(ns) // 20 * 20].reshape(20, -1).mean(1)
    mean, err = ns.mean(), blocks.std(ddof=1) / math.sqrt(20)
    print("two-stage, log theta %.1f: <N> = %.3f +- %.3f (exact %.3f), first stage rejected %s "
          "passed %s, accepted %s" % (log_theta, mean, err, lam, rej, pas, acc))
    assert min(acc) > 1000
    assert abs(mean - lam) < 4 * err + 0.02
    # theta < 1 stops insertions at the first stage (r_1 < 1 when N is near
    # its mean) and lets deletions through; theta > 1 the other way round
    if log_theta < 0:
        assert rej[0] > 0 and pas[0] > 0
    else:
        assert rej[1] > 0 and pas[1] > 0
