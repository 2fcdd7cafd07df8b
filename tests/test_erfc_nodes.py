"""The FP64 erfc of the pair kernels (erfc_nodes in engine_dev.h): Taylor
series about tabulated nodes x_i = i/128 with the Hermite-polynomial
derivatives of erfc (Abramowitz & Stegun 7.1.19), eight terms.

The same steps in double precision are checked here against erfc at 40
digits (mpmath): the method's error over [0, 8) is a few ulp, so the device
real-space terms carry FP64 rounding only.  The device code itself is pinned
end to end by the GPU parity tests (every real-space dE within 1e-10 of the
oracle's libm erfc).  The FP32 pair path's erfc_nodes32 (the same nodes in
float32, four terms) is checked the same way at float32 precision.
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


def _erfc_nodes32(x, nodes32):
    """erfc_nodes32 (engine_dev.h) step by step in float32."""
    f = np.float32
    x = f(x)
    if not x < f(X_MAX):
        return f(0.0)
    i = int(x * f(PER_UNIT) + f(0.5))
    xi = f(i) * f(1.0 / PER_UNIT)
    t = x - xi
    a, e = nodes32[i]
    m2x = f(-2.0) * xi
    q2 = m2x * m2x - f(2.0)
    q3 = m2x * q2 - f(4.0) * m2x
    t2 = t * t
    s = t + t2 * (f(0.5) * m2x + t * (f(1.0 / 6.0) * q2 + t * (f(1.0 / 24.0) * q3)))
    return a - e * s


def test_fp32_erfc_nodes_within_a_few_fp32_ulp_on_0_to_8():
    """The FP32 pair path's erfc: the same nodes rounded to float32, four
    terms of the series, evaluated in float32, against erfc at 40 digits of
    the float32 argument."""
    nodes32 = [(np.float32(a), np.float32(e)) for a, e in _nodes()]
    rng = np.random.default_rng(8)
    xs = np.concatenate([rng.uniform(0, X_MAX, 6000), np.linspace(0, X_MAX - 1e-3, 2001),
                         rng.uniform(0, 1 / 64, 500)]).astype(np.float32)
    worst = 0.0
    for x in xs:
        ref = float(mpmath.erfc(mpmath.mpf(float(x))))
        got = float(_erfc_nodes32(x, nodes32))
        worst = max(worst, abs(got - ref) / ref)
    print("FP32 erfc nodes: worst relative error %.2e (%.1f float32 ulp)" % (worst, worst / 2.0 ** -24))
    assert worst <= 8 * 2.0 ** -24, worst
