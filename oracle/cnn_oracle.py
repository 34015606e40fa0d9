"""CPU oracle of the learned initial guess network (test infrastructure only:
imported by tests/ to check the GPU kernels; never by the product path).

The network has no counterpart in the reference.  It maps five per-cell
features of the radiation solve to the first Newton correction of E
(delta_ER of the first SOLVE_WITH_SOR call, ppenergy.f:370-375):

    x0 = normalised features, 5 channels            (cells of the n_r x n_phi x n_theta grid)
    a1 = gelu(conv(x0; W0, b0))        5 -> 16
    a2 = gelu(conv(a1; W1, b1))       16 -> 16
    a3 = gelu(conv(a2; W2, b2))       16 -> 16
    a4 = gelu(conv(a3; W3, b3))       16 -> 16
    y  =       conv(a4; W4, b4)       16 -> 1

conv is a 3x3x3 "same" convolution over (k, j, i) = (theta, phi, r): zero
padding in r and theta, periodic in phi.  Weights are W[tap][ci][co] with
tap = (dk+1)*9 + (dj+1)*3 + (di+1).  gelu is the tanh form
0.5 u (1 + tanh(sqrt(2/pi) (u + 0.044715 u^3))).  Loss = mean (y - target)^2.

Precision model of the GPU kernels (bf16=True): weights, the stored
activations a1..a4, the layer-0 input and the backward operands dU are
rounded to bfloat16 (round to nearest even); sums, biases, gelu, y and the
loss are float32-or-better.  With bf16=False everything is float64, which
the finite-difference test uses to pin the backward formulas.

Arrays are channels-last over cells ordered (k, j, i): shape (nz, ny, nx, C).
"""
from __future__ import annotations

import numpy as np

CH = [5, 16, 16, 16, 16, 1]
NL = 5
SQ2PI = np.sqrt(2.0 / np.pi)


def n_weights():
    return sum(27 * CH[l] * CH[l + 1] for l in range(NL))


def n_biases():
    return sum(CH[l + 1] for l in range(NL))


def split(w_flat, b_flat):
    """Flat parameter vectors -> per-layer W[tap][ci][co], b[co]."""
    W, B = [], []
    ow = ob = 0
    for l in range(NL):
        ci, co = CH[l], CH[l + 1]
        W.append(np.asarray(w_flat[ow:ow + 27 * ci * co], dtype=np.float64).reshape(27, ci, co))
        B.append(np.asarray(b_flat[ob:ob + co], dtype=np.float64))
        ow += 27 * ci * co
        ob += co
    return W, B


def bf16(a):
    """Round to bfloat16 (nearest even) and back to float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(u):
    return 0.5 * u * (1.0 + np.tanh(SQ2PI * (u + 0.044715 * u ** 3)))


def gelu_d(u):
    t = np.tanh(SQ2PI * (u + 0.044715 * u ** 3))
    return 0.5 * (1.0 + t) + 0.5 * u * (1.0 - t * t) * SQ2PI * (1.0 + 3 * 0.044715 * u * u)


def shift(x, dk, dj, di):
    """s[k, j, i] = x[k+dk, j+dj, i+di]: zero outside in k and i, periodic in j."""
    nz, ny, nx = x.shape[:3]
    s = np.zeros_like(x)
    x = np.roll(x, -dj, axis=1)
    ks = slice(max(0, -dk), min(nz, nz - dk))
    kd = slice(max(0, dk), min(nz, nz + dk))
    is_ = slice(max(0, -di), min(nx, nx - di))
    id_ = slice(max(0, di), min(nx, nx + di))
    s[ks, :, is_] = x[kd, :, id_]
    return s


TAPS = [(dk, dj, di) for dk in (-1, 0, 1) for dj in (-1, 0, 1) for di in (-1, 0, 1)]


def conv(x, W, b):
    y = np.zeros(x.shape[:3] + (W.shape[2],))
    for t, (dk, dj, di) in enumerate(TAPS):
        y += shift(x, dk, dj, di) @ W[t]
    return y + b


def conv_t(g, W):
    """Adjoint of conv's linear part: d/dx of sum(g * conv(x))."""
    dx = np.zeros(g.shape[:3] + (W.shape[1],))
    for t, (dk, dj, di) in enumerate(TAPS):
        dx += shift(g @ W[t].T, -dk, -dj, -di)
    return dx


def forward(x0, w_flat, b_flat, bf=True):
    """Returns y (nz, ny, nx) and the cache for backward."""
    W, B = split(w_flat, b_flat)
    if bf:
        W = [bf16(w) for w in W]
        x0 = bf16(x0)
    acts, pres = [x0], []
    a = x0
    for l in range(NL - 1):
        u = conv(a, W[l], B[l])
        pres.append(u)
        a = gelu(u)
        if bf:
            a = bf16(a)
        acts.append(a)
    y = conv(a, W[NL - 1], B[NL - 1])[..., 0]
    return y, (W, acts, pres, bf)


def loss_grad(x0, target, w_flat, b_flat, bf=True):
    """Mean squared error and its gradients (flat, same layout as the
    parameters)."""
    y, (W, acts, pres, bf_) = forward(x0, w_flat, b_flat, bf)
    n = y.size
    r = y - target
    loss = float(np.mean(r * r))
    g = (2.0 / n) * r[..., None]          # dL/dy, one channel
    gW, gB = [None] * NL, [None] * NL
    for l in range(NL - 1, -1, -1):
        a = acts[l]
        if bf_ and l < NL - 1:
            g = bf16(g)
        gW[l] = np.stack([shift(a, dk, dj, di).reshape(-1, a.shape[3]).T @ g.reshape(-1, g.shape[3])
                          for (dk, dj, di) in TAPS])
        gB[l] = g.reshape(-1, g.shape[3]).sum(axis=0)
        if l == 0:
            break
        da = conv_t(g, W[l])
        g = da * gelu_d(pres[l - 1])
    gw = np.concatenate([x.ravel() for x in gW])
    gb = np.concatenate([x.ravel() for x in gB])
    return loss, gw, gb, y


def init_params(seed=0):
    """He-style initialisation: W ~ N(0, 2 / (27 ci)), b = 0 (the last layer
    scaled down by 10 so the untrained guess starts near zero)."""
    rng = np.random.default_rng(seed)
    ws, bs = [], []
    for l in range(NL):
        ci, co = CH[l], CH[l + 1]
        w = rng.standard_normal((27, ci, co)) * np.sqrt(2.0 / (27 * ci))
        if l == NL - 1:
            w *= 0.1
        ws.append(w.ravel())
        bs.append(np.zeros(co))
    return np.concatenate(ws).astype(np.float32), np.concatenate(bs).astype(np.float32)
