"""CPU checks that pin the float64 surrogate reference (oracle/surrogate.py)
by mathematics: the analytic gradient against central finite differences,
and the descriptor against its invariances."""
import numpy as np

from oracle import surrogate as S


def _net(rng, scale=0.3):
    return rng.normal(0, scale, S.N_PARAMS) / np.sqrt(np.r_[np.full(S.OFF_W2, S.DESCR),
                                                             np.full(S.N_PARAMS - S.OFF_W2, S.HID)])


def test_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    w = _net(rng)
    maxm, samples = 3, 4
    x = rng.normal(0, 1, (samples * 2 * maxm, S.DESCR))
    n = np.array([3, 1, 2, 3])
    target = rng.normal(0, 2, samples)
    target[2] = 50.0          # exercises the clip
    loss, grad = S.loss_and_grad(w, x, target, n, maxm, clip=20.0)
    idx = np.r_[rng.choice(S.OFF_B1, 6, replace=False), S.OFF_B1 + 5, S.OFF_W2 + 777,
                S.OFF_B2 + 3, S.OFF_W3 + 4000, S.OFF_B3 + 9, S.OFF_W4 + 11, S.OFF_B4]
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (S.loss_and_grad(wp, x, target, n, maxm, 20.0)[0] -
              S.loss_and_grad(wm, x, target, n, maxm, 20.0)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])
    # padding rows (m >= n) carry no gradient: changing them changes nothing
    x2 = x.copy()
    x2[1 * 2 * maxm + 2] += 5.0      # sample 1 has n = 1: old row 2 is padding
    assert S.loss_and_grad(w, x2, target, n, maxm, 20.0)[0] == loss


def test_network_output_is_antisymmetric_in_old_new():
    rng = np.random.default_rng(5)
    w = _net(rng)
    a = rng.normal(0, 1, (4, S.DESCR))
    b = rng.normal(0, 1, (4, S.DESCR))
    d_ab = S.forward(w, b).sum() - S.forward(w, a).sum()
    d_ba = S.forward(w, a).sum() - S.forward(w, b).sum()
    assert d_ab == -d_ba


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-5], dtype=np.float32)
    r = S.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie, even mantissa kept
    assert r[2] == 1.0 + 2 ** -7            # above the tie rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-5) <= 3.0e-5 * 2 ** -8


def test_descriptor_invariances():
    rng = np.random.default_rng(9)
    box = [10.0, 10.0, 12.0]
    xyz = rng.uniform(0, 10, (60, 3))
    q = rng.choice([-1.0, 0.0, 1.0], 60)
    p = np.array([5.0, 5.0, 6.0])
    d0 = S.descriptor(p, 1.0, xyz, q, (0, 4), box, 3, 4.0)
    # periodic translation of the neighbours
    shift = xyz + np.array([10.0, -10.0, 12.0])
    assert np.allclose(S.descriptor(p, 1.0, shift, q, (0, 4), box, 3, 4.0), d0, atol=1e-12)
    # permutation of the neighbours outside the excluded range
    perm = np.r_[np.arange(5), 5 + rng.permutation(55)]
    assert np.allclose(S.descriptor(p, 1.0, xyz[perm], q[perm], (0, 4), box, 3, 4.0), d0, atol=1e-12)
    # flipping the bead's charge swaps the like / unlike channels
    d1 = S.descriptor(p, -1.0, xyz, q, (0, 4), box, 3, 4.0)
    assert np.allclose(d1[:16], d0[16:32]) and np.allclose(d1[16:32], d0[:16])
    assert np.allclose(d1[32:48], d0[32:48])
    # a single neighbour at a basis centre: the peak lands in that function
    one = S.descriptor(p, 1.0, np.array([p + [1.7, 0, 0]]), np.array([1.0]), (5, 4), box, 3, 4.0)
    assert np.argmax(one[:16]) == int(round((1.7 - 0.7) / ((4.0 - 0.7) / 15)))
