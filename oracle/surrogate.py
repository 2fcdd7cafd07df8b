"""CPU reference (float64, numpy) of the delayed-acceptance surrogate: the
local descriptors, the 64-128-128-128-1 tanh network, and its training loss
and gradient.  Test infrastructure only: it checks the GPU kernels, it is
never part of the product path.

The surrogate has no counterpart in the reference (the reference evaluates
every trial move exactly), so nothing of the reference pins it.  It is
pinned here by mathematics: the analytic gradient against central finite
differences, and the descriptor against its invariances (translation by a
period, permutation of neighbours, the charge-sign channels).

Layouts follow the device code: parameters W1[64][128], b1, W2[128][128],
b2, W3[128][128], b3, w4[128], b4 (W[in][out]); descriptor = 16 radial basis
functions per channel (q_i q_j > 0, < 0, = 0) scaled by 1/4, then q_i, then
exp(-z/2) and exp(-(Lz - z)/2) with two periodic dimensions, zero padding
to 64.
"""
import math

import numpy as np

DESCR, HID, RBF, CHAN = 64, 128, 16, 3
RBF_MIN, FEAT_SCALE = 0.7, 0.25
OFF_W1 = 0
OFF_B1 = OFF_W1 + DESCR * HID
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID * HID
OFF_W4 = OFF_B3 + HID
OFF_B4 = OFF_W4 + HID
N_PARAMS = OFF_B4 + 1


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return (w[OFF_W1:OFF_B1].reshape(DESCR, HID), w[OFF_B1:OFF_W2],
            w[OFF_W2:OFF_B2].reshape(HID, HID), w[OFF_B2:OFF_W3],
            w[OFF_W3:OFF_B3].reshape(HID, HID), w[OFF_B3:OFF_W4],
            w[OFF_W4:OFF_B4], w[OFF_B4])


def forward(w, x, round_bf16=False):
    """f(x) for rows x (M, 64).  round_bf16 mimics the inference kernel: inputs,
    weights and hidden activations rounded to BF16 (nearest even), FP32-like
    accumulation done here in float64."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(w)
    r = bf16 if round_bf16 else (lambda a: a)
    h1 = r(np.tanh(r(np.asarray(x, dtype=np.float64)) @ r(W1) + b1))
    h2 = r(np.tanh(h1 @ r(W2) + b2))
    h3 = r(np.tanh(h2 @ r(W3) + b3))
    return h3 @ w4 + b4


def bf16(a):
    """Round float32 values to BF16 (round to nearest even), returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def loss_and_grad(w, x, target, n, maxm, clip):
    """Mean squared error of sum_new f - sum_old f against clip(target) over
    samples (sample i: rows [2 i maxm, 2 (i + 1) maxm), n[i] old beads then
    n[i] new beads from row maxm on), and its gradient (backpropagation)."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(w)
    x = np.asarray(x, dtype=np.float64)
    S = len(target)
    z1 = x @ W1 + b1
    h1 = np.tanh(z1)
    h2 = np.tanh(h1 @ W2 + b2)
    h3 = np.tanh(h2 @ W3 + b3)
    f = h3 @ w4 + b4
    g = np.zeros(len(x))
    loss = 0.0
    for i in range(S):
        o = i * 2 * maxm
        pred = f[o + maxm:o + maxm + n[i]].sum() - f[o:o + n[i]].sum()
        res = pred - min(max(target[i], -clip), clip)
        loss += res * res
        g[o:o + n[i]] = -2 * res / S
        g[o + maxm:o + maxm + n[i]] = 2 * res / S
    loss /= S
    grad = np.zeros(N_PARAMS)
    grad[OFF_W4:OFF_B4] = h3.T @ g
    grad[OFF_B4] = g.sum()
    d3 = np.outer(g, w4) * (1 - h3 ** 2)
    grad[OFF_W3:OFF_B3] = (h2.T @ d3).ravel()
    grad[OFF_B3:OFF_W4] = d3.sum(0)
    d2 = (d3 @ W3.T) * (1 - h2 ** 2)
    grad[OFF_W2:OFF_B2] = (h1.T @ d2).ravel()
    grad[OFF_B2:OFF_W3] = d2.sum(0)
    d1 = (d2 @ W2.T) * (1 - h1 ** 2)
    grad[OFF_W1:OFF_B1] = (x.T @ d1).ravel()
    grad[OFF_B1:OFF_W2] = d1.sum(0)
    return loss, grad


def descriptor(p, qi, xyz, q, exclude, box, npbc, r_d):
    """Descriptor of a bead of charge qi at p against beads xyz (N, 3), q (N,)
    other than the index range `exclude` = (lo, hi), minimum image in the
    periodic dimensions (x, y; z too when npbc == 3)."""
    xyz = np.asarray(xyz, dtype=np.float64)
    d = xyz - np.asarray(p, dtype=np.float64)
    for k in range(3):
        if k < 2 or npbc == 3:
            d[:, k] -= box[k] * np.round(d[:, k] / box[k])
    r = np.sqrt((d * d).sum(1))
    keep = r < r_d
    lo, hi = exclude
    idx = np.arange(len(xyz))
    keep &= ~((idx >= lo) & (idx <= hi))
    r, qj = r[keep], np.asarray(q)[keep]
    dmu = (r_d - RBF_MIN) / (RBF - 1)
    gam = 1.0 / (dmu * dmu)
    mu = RBF_MIN + dmu * np.arange(RBF)
    fc = 0.5 * (np.cos(math.pi * r / r_d) + 1.0)
    phi = np.exp(-gam * (r[:, None] - mu[None, :]) ** 2) * fc[:, None]
    out = np.zeros(DESCR)
    qq = qi * qj
    for c, sel in enumerate((qq > 0, qq < 0, qq == 0)):
        out[c * RBF:(c + 1) * RBF] = FEAT_SCALE * phi[sel].sum(0)
    out[RBF * CHAN] = qi
    if npbc == 2:
        out[RBF * CHAN + 1] = math.exp(-0.5 * p[2])
        out[RBF * CHAN + 2] = math.exp(-0.5 * (box[2] - p[2]))
    return out
