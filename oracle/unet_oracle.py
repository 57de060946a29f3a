"""FP64 CPU restatement of the learned Poisson preconditioner network.

TEST INFRASTRUCTURE ONLY: imported by tests/ and never by the product path.

The network is the linear three-level 3-D convolutional U-Net of
cyclix_msparc_amd/csrc/cyx_unet.hip (it has no counterpart in the reference,
whose Poisson preconditioner is ILU, src/poissonSolve.m:36).  Tensors of a
level are (nz, nt, nr, C) arrays (r fastest in memory, channels last).
Parameters are one flat vector: per layer W[tap][cout][cin], then the
per-radius output scale s[nr].

conv3: 3x3x3, zero padding in r, periodic in theta and z, tap
(dz+1)*9 + (dt+1)*3 + (dr+1).  down: 2x2x2 stride 2, tap c*4 + b*2 + a for
fine offsets (z, theta, r) = (c, b, a).  up: the transpose of down.
"""
from __future__ import annotations

import numpy as np

CONV3, DOWN, UP = 0, 1, 2
# e0 d1 e1 d2 e2 m2 u1 f1 u0 f0: (mode, cin, cout)
SPEC = [(CONV3, 1, 16), (DOWN, 16, 32), (CONV3, 32, 32), (DOWN, 32, 64), (CONV3, 64, 64),
        (CONV3, 64, 64), (UP, 64, 32), (CONV3, 32, 32), (UP, 32, 16), (CONV3, 16, 1)]


def ntap(mode):
    return 27 if mode == CONV3 else 8


def offsets(nr):
    off = [0]
    for mode, ci, co in SPEC:
        off.append(off[-1] + ntap(mode) * ci * co)
    return off, off[-1] + nr


def unpack(params, nr):
    off, n = offsets(nr)
    assert params.size == n
    Ws = [params[off[l]:off[l + 1]].reshape(ntap(SPEC[l][0]), SPEC[l][2], SPEC[l][1]) for l in range(len(SPEC))]
    return Ws, params[off[-1]:]


def _shift(x, dz, dt, dr):
    """y[k, j, i] = x[k+dz, j+dt, i+dr] with z, theta periodic and zero outside r."""
    y = np.roll(np.roll(x, -dz, axis=0), -dt, axis=1)
    out = np.zeros_like(y)
    nr = x.shape[2]
    if dr >= 0:
        out[:, :, :nr - dr] = y[:, :, dr:]
    else:
        out[:, :, -dr:] = y[:, :, :nr + dr]
    return out


def conv3(x, W):
    out = np.zeros(x.shape[:3] + (W.shape[1],))
    for tap in range(27):
        dz, dt, dr = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        out += _shift(x, dz, dt, dr) @ W[tap].T
    return out


def conv3_adj(g, W):
    """Adjoint of conv3 in its input."""
    out = np.zeros(g.shape[:3] + (W.shape[2],))
    for tap in range(27):
        dz, dt, dr = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        out += _shift(g @ W[tap], -dz, -dt, -dr)
    return out


def down(x, W):
    nz, nt, nr, _ = x.shape
    out = np.zeros((nz // 2, nt // 2, nr // 2, W.shape[1]))
    for tap in range(8):
        c, b, a = tap >> 2, (tap >> 1) & 1, tap & 1
        out += x[c::2, b::2, a::2] @ W[tap].T
    return out


def up(x, W):
    nz, nt, nr, _ = x.shape
    out = np.zeros((2 * nz, 2 * nt, 2 * nr, W.shape[1]))
    for tap in range(8):
        c, b, a = tap >> 2, (tap >> 1) & 1, tap & 1
        out[c::2, b::2, a::2] = x @ W[tap].T
    return out


def down_adj(g, W):
    nz, nt, nr, _ = g.shape
    out = np.zeros((2 * nz, 2 * nt, 2 * nr, W.shape[2]))
    for tap in range(8):
        c, b, a = tap >> 2, (tap >> 1) & 1, tap & 1
        out[c::2, b::2, a::2] = g @ W[tap]
    return out


def up_adj(g, W):
    nz, nt, nr, _ = g.shape
    out = np.zeros((nz // 2, nt // 2, nr // 2, W.shape[2]))
    for tap in range(8):
        c, b, a = tap >> 2, (tap >> 1) & 1, tap & 1
        out += g[c::2, b::2, a::2] @ W[tap]
    return out


def forward(params, x, dims):
    """x: (N,) in r-fastest order; returns (U(x) as (N,), cache)."""
    nr, nt, nz = dims
    Ws, s = unpack(params, nr)
    X = x.reshape(nz, nt, nr, 1)
    a0 = conv3(X, Ws[0])
    a1 = down(a0, Ws[1])
    b1 = conv3(a1, Ws[2])
    a2 = down(b1, Ws[3])
    b2 = conv3(a2, Ws[4])
    c2 = conv3(b2, Ws[5])
    c1 = up(c2, Ws[6]) + b1
    g1 = conv3(c1, Ws[7])
    c0 = up(g1, Ws[8]) + a0
    y = conv3(c0, Ws[9])[..., 0]
    u = y * s[None, None, :]
    cache = dict(X=X, a0=a0, a1=a1, b1=b1, a2=a2, b2=b2, c2=c2, c1=c1, g1=g1, c0=c0, y=y)
    return u.reshape(-1), cache


def _dw_conv3(gout, x):
    dW = np.zeros((27, gout.shape[-1], x.shape[-1]))
    for tap in range(27):
        dz, dt, dr = tap // 9 - 1, (tap // 3) % 3 - 1, tap % 3 - 1
        xs = _shift(x, dz, dt, dr)
        dW[tap] = gout.reshape(-1, gout.shape[-1]).T @ xs.reshape(-1, xs.shape[-1])
    return dW


def _dw_down(gout, x):
    dW = np.zeros((8, gout.shape[-1], x.shape[-1]))
    for tap in range(8):
        c, b, a = tap >> 2, (tap >> 1) & 1, tap & 1
        xs = x[c::2, b::2, a::2]
        dW[tap] = gout.reshape(-1, gout.shape[-1]).T @ xs.reshape(-1, xs.shape[-1])
    return dW


def _dw_up(gout, x):
    dW = np.zeros((8, gout.shape[-1], x.shape[-1]))
    for tap in range(8):
        c, b, a = tap >> 2, (tap >> 1) & 1, tap & 1
        gs = gout[c::2, b::2, a::2]
        dW[tap] = gs.reshape(-1, gs.shape[-1]).T @ x.reshape(-1, x.shape[-1])
    return dW


def loss_and_grad(params, X, T, dims):
    """L = (1/nb) sum_b ||U(x_b) - t_b||^2 / ||t_b||^2 and dL/dparams."""
    nr, nt, nz = dims
    Ws, s = unpack(params, nr)
    nb = X.shape[0]
    grads = [np.zeros_like(W) for W in Ws]
    gs = np.zeros_like(s)
    L = 0.0
    for b in range(nb):
        u, c = forward(params, X[b], dims)
        w = 1.0 / (nb * np.dot(T[b], T[b]))
        e = (u - T[b]).reshape(nz, nt, nr)
        L += w * np.sum(e * e)
        du = 2 * w * e
        gs += np.sum(du * c["y"], axis=(0, 1))
        dy = (du * s[None, None, :])[..., None]
        grads[9] += _dw_conv3(dy, c["c0"])
        dc0 = conv3_adj(dy, Ws[9])
        grads[8] += _dw_up(dc0, c["g1"])
        dg1 = up_adj(dc0, Ws[8])
        grads[7] += _dw_conv3(dg1, c["c1"])
        dc1 = conv3_adj(dg1, Ws[7])
        grads[6] += _dw_up(dc1, c["c2"])
        dc2 = up_adj(dc1, Ws[6])
        grads[5] += _dw_conv3(dc2, c["b2"])
        db2 = conv3_adj(dc2, Ws[5])
        grads[4] += _dw_conv3(db2, c["a2"])
        da2 = conv3_adj(db2, Ws[4])
        grads[3] += _dw_down(da2, c["b1"])
        db1 = dc1 + down_adj(da2, Ws[3])
        grads[2] += _dw_conv3(db1, c["a1"])
        da1 = conv3_adj(db1, Ws[2])
        grads[1] += _dw_down(da1, c["a0"])
        da0 = dc0 + down_adj(da1, Ws[1])
        grads[0] += _dw_conv3(da0, c["X"])
    g = np.concatenate([gw.reshape(-1) for gw in grads] + [gs])
    return L, g


def bf16_round(a):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)
