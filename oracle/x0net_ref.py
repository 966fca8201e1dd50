"""NumPy restatement of the initial-guess network's arithmetic.

TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels of
libemmi_mgfd_amd/csrc/x0net_kernels.hip; nothing in the product imports it.

Layouts follow the kernels: tensors [Z][Y][X][C], weights [Cout][27*Cin] with
tap = 9*dz + 3*dy + dx (dz, dy, dx in 0..2, offset -1), zero padding 1.
All arithmetic in float64.
"""
from __future__ import annotations

import numpy as np

SLOPE = 0.1


def to_bf16_bits(a):
    """float -> bf16 bit patterns (uint16), round to nearest even."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16_round(a):
    return from_bf16_bits(to_bf16_bits(a)).astype(np.float64)


def taps():
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                yield 9 * dz + 3 * dy + dx, dz - 1, dy - 1, dx - 1


def shifted(x, dz, dy, dx):
    """s[z, y, x] = x[z+dz, y+dy, x+dx] with zeros outside."""
    Z, Y, X = x.shape[:3]
    p = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    return p[1 + dz:1 + dz + Z, 1 + dy:1 + dy + Y, 1 + dx:1 + dx + X]


def conv_fwd(x, w, b=None):
    Cin = x.shape[3]
    Cout = w.shape[0]
    y = np.zeros(x.shape[:3] + (Cout,))
    for t, dz, dy, dx in taps():
        y += shifted(x, dz, dy, dx) @ w[:, t * Cin:(t + 1) * Cin].T
    if b is not None:
        y += b
    return y


def conv_dx(dy_, w, cin):
    dx = np.zeros(dy_.shape[:3] + (cin,))
    for t, dz, dyy, dxx in taps():
        # x[v + off] contributed to y[v]; so dx[u] += dy[u - off] w_t
        dx += shifted(dy_, -dz, -dyy, -dxx) @ w[:, t * cin:(t + 1) * cin]
    return dx


def conv_dw(x, dy_):
    Cin = x.shape[3]
    Cout = dy_.shape[3]
    dw = np.zeros((Cout, 27 * Cin))
    D = dy_.reshape(-1, Cout)
    for t, dz, dyy, dxx in taps():
        dw[:, t * Cin:(t + 1) * Cin] = D.T @ shifted(x, dz, dyy, dxx).reshape(-1, Cin)
    return dw


def leaky(z):
    return np.where(z > 0, z, SLOPE * z)


def dleaky_from_out(a):
    return np.where(a > 0, 1.0, SLOPE)


def pool2(x):
    Z, Y, X, C = x.shape
    return x.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).mean(axis=(1, 3, 5))


def pool2_bwd(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def up2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up2_bwd(g):
    Z, Y, X, C = g.shape
    return g.reshape(Z // 2, 2, Y // 2, 2, X // 2, 2, C).sum(axis=(1, 3, 5))


def _ident(a):
    return a


def unet_forward(feat, params, levels, rnd=_ident):
    """params: list of (w, b) in the order encoder (2 per level), decoder
    (2 per level, coarsest first), output.  rnd is applied wherever the
    kernels store a tensor in BF16 (bf16_round to emulate them).  Returns the
    output and a tape."""
    nl = len(levels)
    tape = {"enc": [], "pool": [], "cat": [], "dec": []}
    it = iter(params)
    x = feat
    for lev in range(nl):
        (w1, b1), (w2, b2) = next(it), next(it)
        a1 = rnd(leaky(conv_fwd(x, w1, b1)))
        a2 = rnd(leaky(conv_fwd(a1, w2, b2)))
        tape["enc"].append((x, a1, a2))
        if lev < nl - 1:
            x = rnd(pool2(a2))
            tape["pool"].append(x)
    up = tape["enc"][-1][2]
    for i, lev in enumerate(range(nl - 2, -1, -1)):
        cat = np.concatenate([up2(up), tape["enc"][lev][2]], axis=3)
        (w1, b1), (w2, b2) = next(it), next(it)
        a1 = rnd(leaky(conv_fwd(cat, w1, b1)))
        a2 = rnd(leaky(conv_fwd(a1, w2, b2)))
        tape["cat"].append(cat)
        tape["dec"].append((a1, a2))
        up = a2
    wo, bo = next(it)
    y = conv_fwd(up, wo, bo)
    tape["top"] = up
    return y, tape


def unet_backward(gy, params, levels, tape, rnd=_ident):
    """Parameter gradients (same order as params) for output gradient gy,
    with rnd applied where the kernels store BF16 gradients."""
    nl = len(levels)
    grads = [None] * len(params)
    ne = 2 * nl
    wo, _ = params[-1]
    grads[-1] = (conv_dw(tape["top"], gy), gy.sum(axis=(0, 1, 2)))
    g = rnd(conv_dx(gy, wo, wo.shape[1] // 27))
    gskip = [None] * nl
    for i in range(nl - 2, -1, -1):
        lev = nl - 2 - i
        (w1, _), (w2, _) = params[ne + 2 * i], params[ne + 2 * i + 1]
        a1, a2 = tape["dec"][i]
        dz2 = rnd(g * dleaky_from_out(a2))
        grads[ne + 2 * i + 1] = (conv_dw(a1, dz2), dz2.sum(axis=(0, 1, 2)))
        dz1 = rnd(rnd(conv_dx(dz2, w2, w2.shape[1] // 27)) * dleaky_from_out(a1))
        cat = tape["cat"][i]
        grads[ne + 2 * i] = (conv_dw(cat, dz1), dz1.sum(axis=(0, 1, 2)))
        gcat = rnd(conv_dx(dz1, w1, cat.shape[3]))
        c_up = levels[lev + 1]
        gskip[lev] = gcat[..., c_up:]
        g = rnd(up2_bwd(gcat[..., :c_up]))
    for lev in range(nl - 1, -1, -1):
        (w1, _), (w2, _) = params[2 * lev], params[2 * lev + 1]
        x, a1, a2 = tape["enc"][lev]
        dz2 = rnd(g * dleaky_from_out(a2))
        grads[2 * lev + 1] = (conv_dw(a1, dz2), dz2.sum(axis=(0, 1, 2)))
        dz1 = rnd(rnd(conv_dx(dz2, w2, w2.shape[1] // 27)) * dleaky_from_out(a1))
        grads[2 * lev] = (conv_dw(x, dz1), dz1.sum(axis=(0, 1, 2)))
        if lev > 0:
            gin = rnd(conv_dx(dz1, w1, x.shape[3]))
            g = rnd(pool2_bwd(gin) + gskip[lev - 1])
    return grads
