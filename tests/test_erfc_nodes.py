"""The FP64 erfc of the pair kernels (erfc_nodes in engine_dev.h): Taylor
series about tabulated nodes x_i = i/128 with the Hermite-polynomial
derivatives of erfc (Abramowitz & Stegun 7.1.19), eight terms.

The same steps in double precision are checked here against erfc at 40
digits (mpmath): the method's error over [0, 8) is a few ulp, so the device
real-space terms carry FP64 rounding only.  The device code itself is pinned
end to end by the GPU parity tests (every real-space dE within 1e-10 of the
oracle's libm erfc).
"""
import math

import numpy as np
import pytest

mpmath = pytest.importorskip("mpmath")

PER_UNIT = 128
X_MAX = 8.0


def _nodes():
    mpmath.mp.dps = 40
    out = []
    for i in range(int(X_MAX) * PER_UNIT + 1):
        x = mpmath.mpf(i) / PER_UNIT
        out.append((float(mpmath.erfc(x)), float(2 / mpmath.sqrt(mpmath.pi) * mpmath.exp(-x * x))))
    return out


def _erfc_nodes(x, nodes):
    if not x < X_MAX:
        return 0.0
    i = int(x * PER_UNIT + 0.5)
    xi = i * (1.0 / PER_UNIT)
    t = x - xi
    a, e = nodes[i]
    m2x = -2.0 * xi
    qm, q = 1.0, m2x
    w = t
    s = w
    w *= t * 0.5
    s += q * w
    for n in range(1, 7):
        qn = m2x * q - (2.0 * n) * qm
        qm, q = q, qn
        w *= t * (1.0 / (n + 2))
        s += q * w
    return a - e * s


def test_erfc_nodes_within_four_ulp_on_0_to_8():
    nodes = _nodes()
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.uniform(0, X_MAX, 6000), np.linspace(0, X_MAX - 1e-3, 2001),
                         rng.uniform(0, 1 / 64, 500), [0.0, 1 / 256, 0.5 - 1 / 256, 3.99609375]])
    worst = 0.0
    for x in xs:
        ref = float(mpmath.erfc(mpmath.mpf(float(x))))
        got = _erfc_nodes(float(x), nodes)
        worst = max(worst, abs(got - ref) / ref)
    assert worst <= 4 * 2.0 ** -53, worst


def test_erfc_nodes_exact_at_the_nodes_and_zero_beyond_range():
    nodes = _nodes()
    for i in (0, 1, 64, 255, 700, 1023):
        x = i / PER_UNIT
        assert _erfc_nodes(x, nodes) == nodes[i][0]
    assert _erfc_nodes(8.0, nodes) == 0.0
    assert math.erfc(8.0) < 1.2e-29
