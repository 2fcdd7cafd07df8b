"""Pins of the oracle pieces that no reference fixture covers: the harmonic
bond-length sampler of the configurational-bias growth and the branches of
the 6-3 wall for right-grafted ("R") beads and for a wall cutoff >= 0.

Each check is chosen so that a plausible slip in the restatement fails it:
the sampler against the density its acceptance rule fixes (a len/a rule or a
missing thinning moves the mean by many standard errors), the "R" branch
against the mirror image of the "L" branch and a closed form, the cut wall
against its closed form, continuity and well depth."""
import math

import numpy as np
import pytest

from oracle import oracle as O

A = 3 * math.sqrt(3) / 2      # the 6-3 wall prefactor, 2.59807621135


# ------------------------------------------------ RandomBondLen sampler
def _spring_sys(k_bond, r0, beta):
    sd = dict(box=[10, 10, 10], npbc=3, beta=beta, use_pair=0, use_ewald=0, use_bond=1,
              use_ext=0, k_bond=k_bond, r0=r0,
              types=[{"name": "P", "sigma": 1.0, "eps": 1.0}],
              xyz=np.array([[1.0, 1.0, 1.0], [2.0, 1.0, 1.0]]), q=np.zeros(2),
              type=np.zeros(2, dtype=np.int32), mol_start=np.array([0, 2], dtype=np.int32))
    return O.Oracle(sd)


def _density(r0, sigma, weight):
    """Grid, normalised density and CDF of N(r0, sigma) times `weight`."""
    r = np.linspace(r0 - 9 * sigma, r0 + 9 * sigma, 200001)
    p = np.exp(-0.5 * ((r - r0) / sigma) ** 2) * weight(r)
    p /= np.trapezoid(p, r)
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (p[1:] + p[:-1]) * np.diff(r))])
    return r, p, cdf / cdf[-1]


def _moments(r, p):
    m = np.trapezoid(r * p, r)
    return m, np.trapezoid((r - m) ** 2 * p, r)


@pytest.mark.parametrize("k_bond,r0,beta", [(400.0, 1.0, 1.0), (25.0, 1.0, 1.0), (10.0, 1.5, 2.0)])
def test_random_bond_len_samples_the_thinned_gaussian(k_bond, r0, beta):
    """potential_spring.cc:50-61: a Gaussian draw around r0 of width
    1/sqrt(beta k) is kept with probability len^2 / (r0 + 3 sigma)^2, so the
    lengths follow N(r0, sigma) x min(r^2, (r0 + 3 sigma)^2)."""
    o = _spring_sys(k_bond, r0, beta)
    sigma = 1.0 / math.sqrt(beta * k_bond)
    a = (r0 + 3 * sigma) ** 2
    n = 200000
    x = o.random_bond_lengths(O.MT19937(1234), O.OrcGauss(), n)
    o.close()
    r, p, cdf = _density(r0, sigma, lambda t: np.minimum(t * t, a))
    mean, var = _moments(r, p)
    se_m = math.sqrt(var / n)
    se_v = var * math.sqrt(2.0 / n)
    assert abs(x.mean() - mean) < 5 * se_m, (x.mean(), mean, se_m)
    assert abs(x.var() - var) < 5 * se_v, (x.var(), var, se_v)
    # chi^2 over 40 equiprobable bins of the target distribution
    edges = np.interp(np.linspace(0, 1, 41), cdf, r)
    edges[0], edges[-1] = -np.inf, np.inf
    obs = np.histogram(x, bins=edges)[0]
    chi2 = float(((obs - n / 40.0) ** 2 / (n / 40.0)).sum())
    assert chi2 < 39 + 6 * math.sqrt(2 * 39), chi2
    # the check has power against the plausible slips: a len / sqrt(a)
    # acceptance or no thinning at all moves the mean by many errors
    for wrong in (lambda t: np.minimum(np.abs(t), math.sqrt(a)), lambda t: np.ones_like(t)):
        m_wrong, _ = _moments(*_density(r0, sigma, wrong)[:2])
        assert abs(x.mean() - m_wrong) > 8 * se_m, (x.mean(), m_wrong, se_m)


# ------------------------------------------------ 6-3 wall, branches
def _wall(z, kind="P", wall_cut=-1.0, Lz=10.0, sigma=1.0, eps=2.0):
    types = [{"name": kind, "sigma": 1.0, "eps": 1.0, "wall_sigma": sigma, "wall_eps": eps}]
    sd = dict(box=[10, 10, Lz], npbc=2, beta=1, use_pair=0, use_ewald=0, use_bond=0,
              use_ext=1, types=types, xyz=np.array([[1.0, 1.0, z]]), q=np.zeros(1),
              type=np.zeros(1, dtype=np.int32), mol_start=np.array([0, 1], dtype=np.int32),
              wall_cut=wall_cut)
    o = O.Oracle(sd)
    e = o.wall(0)
    o.close()
    return e


@pytest.mark.parametrize("z", [0.3, 0.9, 1.2, 1.5, 2.0, 3.0, 3.7])
@pytest.mark.parametrize("sigma", [1.0, 1.3])
def test_wall_right_grafted_mirrors_left_grafted(z, sigma):
    """An "R" bead at distance d from the z = Lz wall sees what an "L" bead
    at distance d from the z = 0 wall sees (well + FENE tether at its own
    wall, repulsive or constant far wall)."""
    Lz = 10.0
    assert _wall(Lz - z, "R", sigma=sigma) == pytest.approx(_wall(z, "L", sigma=sigma), rel=1e-12)
    # a short box: the far wall inside its repulsive range for both
    Lz = 3.0
    if z < Lz:
        assert _wall(Lz - z, "R", Lz=Lz, sigma=sigma) == pytest.approx(
            _wall(z, "L", Lz=Lz, sigma=sigma), rel=1e-12)


def test_wall_right_grafted_closed_form():
    eps, Lz = 2.0, 3.0
    R0 = 3 * 2 ** (1 / 3)
    # d = 1.2 from the right wall; the left wall at 1.8 > 2^(1/3): its
    # constant far-wall value
    z = Lz - 1.2
    d = 1.2
    fene = -0.5 * R0 * R0 * math.log(1 - (d / R0) ** 2)
    near = A * eps * ((1 / d) ** 6 - (1 / d) ** 3)
    far = A * eps * (0.25 - 0.5)
    assert _wall(z, "R", Lz=Lz, eps=eps) == pytest.approx(near + fene + far, rel=1e-9)
    # d = 2.0 from the right wall: the left wall at 1.0 < 2^(1/3), unshifted
    z = 1.0
    d = Lz - z
    fene = -0.5 * R0 * R0 * math.log(1 - (d / R0) ** 2)
    near = A * eps * ((1 / d) ** 6 - (1 / d) ** 3)
    far = A * eps * (1.0 - 1.0)
    # (the reference's A and 2^(1/3) are rounded to 12 digits: rel 1e-9)
    assert _wall(z, "R", Lz=Lz, eps=eps) == pytest.approx(near + fene + far, rel=1e-9)
    # on or beyond either wall: the 1e8 sentinel
    assert _wall(0.0, "R", Lz=Lz) == 1e8 and _wall(Lz, "R", Lz=Lz) == 1e8


@pytest.mark.parametrize("cut", [1.5, 2.5])
def test_wall_with_cutoff_closed_form(cut):
    """wall_cut >= 0 (potential_truncated_lj_wall.cc:119-130): the full 6-3
    well at each wall within `cut`, shifted by its value at cut for sigma = 1
    (the shift uses 1/cut, not sigma/cut)."""
    eps, Lz = 2.0, 10.0
    shift = A * eps * ((1 / cut) ** 6 - (1 / cut) ** 3)
    for z in (0.7, 1.0, 2 ** (1 / 3), 1.4):
        if z >= cut:
            continue
        e = A * eps * ((1 / z) ** 6 - (1 / z) ** 3) - shift
        assert _wall(z, wall_cut=cut, eps=eps) == pytest.approx(e, rel=1e-9)
        assert _wall(Lz - z, wall_cut=cut, eps=eps) == pytest.approx(_wall(z, wall_cut=cut, eps=eps), rel=1e-12)
    # the well bottom: -A eps / 4 at z = 2^(1/3) sigma
    zm = 2 ** (1 / 3)
    assert _wall(zm, wall_cut=cut, eps=eps) == pytest.approx(-A * eps / 4 - shift, rel=1e-9)
    # continuous (zero) at the cut for sigma = 1, zero beyond it
    assert abs(_wall(cut - 1e-9, wall_cut=cut, eps=eps)) < 1e-7
    assert _wall(cut + 1e-9, wall_cut=cut, eps=eps) == 0.0
    # sigma != 1: the shift stays at 1/cut, so the wall jumps at the cut by
    # A eps [(s/cut)^6 - (s/cut)^3 - (1/cut)^6 + (1/cut)^3]
    s = 1.2
    jump = A * eps * ((s / cut) ** 6 - (s / cut) ** 3) - shift
    assert _wall(cut - 1e-12, wall_cut=cut, eps=eps, sigma=s) == pytest.approx(jump, rel=1e-6)
    # both walls within the cut in a thin slit add up
    Lz2 = 1.9 if cut > 1.0 else 1.2
    z = 0.8
    both = (A * eps * ((1 / z) ** 6 - (1 / z) ** 3) - shift +
            A * eps * ((1 / (Lz2 - z)) ** 6 - (1 / (Lz2 - z)) ** 3) - shift)
    assert _wall(z, wall_cut=cut, eps=eps, Lz=Lz2) == pytest.approx(both, rel=1e-9)
