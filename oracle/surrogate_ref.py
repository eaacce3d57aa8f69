"""FP64 restatement of the learned warm start (test infrastructure only).

The surrogate has no counterpart in the reference (its GMRES starts from
zero, gmres.rs:38-39); this module is the mathematical definition the HIP
kernels are checked against:
  * features of a collocation point under a plane wave (32 per point) and
    the training targets (psi = phi / phi_inc at unit RMS per system),
  * the network 32 -> W -> W -> W -> 2 with SiLU, its MSE loss and the
    analytic gradient (pinned by central finite differences in the tests),
  * one Adam step (decoupled weight decay on weights only).
With ``emulate=True`` every value the HIP path stores in BF16 (features,
W1..W3, hidden pre-activations and activations, back-propagated dZ) is
rounded to BF16 at the same point, so the GPU must match to FP32 rounding.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use this.
"""
from __future__ import annotations

import numpy as np

FEAT = 32


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    out = u.astype(np.uint32).view(np.float32).astype(np.float64)
    if np.any(nan):
        out = np.where(nan, np.nan, out)
    return out


def bf16_bits_to_float(h):
    return (np.asarray(h, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def system_params(cpts, frequency, sound_speed, mass_density, direction, amplitude):
    """k, unit direction, A/(i omega rho), centroid and RMS radius of the mesh."""
    omega = 2.0 * np.pi * frequency
    k = omega / sound_speed
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    amp = complex(amplitude[0], amplitude[1]) / (1j * omega * mass_density)
    xyz = cpts[:, 0:3]
    c0 = xyz.mean(axis=0)
    a = float(np.sqrt(np.mean(np.sum((xyz - c0) ** 2, axis=1))))
    return {"k": k, "d": d, "amp": amp, "c0": c0, "a": a}


def features(cpts, sp):
    """(n, 32) features (FP64, not rounded) of one system sp on the mesh."""
    xyz, nrm = cpts[:, 0:3], cpts[:, 3:6]
    r = xyz - sp["c0"]
    rl = np.linalg.norm(r, axis=1)
    h = r / np.where(rl > 0, rl, 1.0)[:, None] * (rl > 0)[:, None]
    d, k, a = sp["d"], sp["k"], sp["a"]
    ka = k * a
    dr = h @ d
    dn = nrm @ d
    nr = np.sum(nrm * h, axis=1)
    th = np.arccos(np.clip(dr, -1.0, 1.0))
    dx = r @ d
    f = np.zeros((xyz.shape[0], FEAT))
    f[:, 0] = ka / 8.0
    f[:, 1] = (ka / 8.0) ** 2
    f[:, 2] = dr
    f[:, 3] = dn
    f[:, 4] = nr
    f[:, 5] = rl / a - 1.0
    f[:, 6:9] = h
    f[:, 9:12] = nrm
    f[:, 12] = th / np.pi
    f[:, 13] = np.cos(ka * dr)
    f[:, 14] = np.sin(ka * dr)
    f[:, 15] = np.cos(ka * th)
    f[:, 16] = np.sin(ka * th)
    f[:, 17] = np.cos(ka * (np.pi - th))
    f[:, 18] = np.sin(ka * (np.pi - th))
    f[:, 19] = np.cos(2.0 * ka * dr)
    f[:, 20] = np.sin(2.0 * ka * dr)
    f[:, 21] = np.cos(ka)
    f[:, 22] = np.sin(ka)
    f[:, 23] = np.cos(k * (rl - dx))
    f[:, 24] = np.sin(k * (rl - dx))
    f[:, 25] = 1.0 / (1.0 + ka)
    f[:, 26] = cpts[:, 6] / cpts[:, 6].mean() - 1.0
    return f


def incident(cpts, sp):
    return sp["amp"] * np.exp(1j * sp["k"] * (cpts[:, 0:3] @ sp["d"]))


def targets(cpts, sp, phi, normalize=True):
    """psi = phi / phi_inc as (n, 2) (re, im), divided by its RMS over the
    system's points when normalize."""
    psi = np.asarray(phi) / incident(cpts, sp)
    if normalize:
        psi = psi / np.sqrt(np.mean(np.abs(psi) ** 2))
    return np.stack([psi.real, psi.imag], axis=1)


def layout(W):
    """Offsets (w, b) of the four layers in the flat parameter vector."""
    fan_in, fan_out = [FEAT, W, W, W], [W, W, W, 2]
    offs, o = [], 0
    for l in range(4):
        ow = o
        o += fan_out[l] * fan_in[l]
        ob = o
        o += fan_out[l]
        offs.append((ow, ob, fan_out[l], fan_in[l]))
    return offs, o


def unpack(params, W):
    offs, total = layout(W)
    assert params.size == total
    Ws, bs = [], []
    for ow, ob, fo, fi in offs:
        Ws.append(np.asarray(params[ow:ow + fo * fi], dtype=np.float64).reshape(fo, fi))
        bs.append(np.asarray(params[ob:ob + fo], dtype=np.float64))
    return Ws, bs


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, X, W, emulate=False):
    r = bf16 if emulate else (lambda v: v)
    Ws, bs = unpack(params, W)
    a = r(X)
    cache = {"A": [a], "Z": []}
    for l in range(3):
        z = a @ r(Ws[l]).T + bs[l]
        cache["Z"].append(r(z))
        a = r(silu(z))
        cache["A"].append(a)
    out = a @ Ws[3].T + bs[3]
    return out, cache


def loss_grad(params, X, T, W, emulate=False):
    """Mean over points of |out - T|^2 and its gradient (flat, like params)."""
    r = bf16 if emulate else (lambda v: v)
    Ws, _ = unpack(params, W)
    out, c = forward(params, X, W, emulate)
    P = X.shape[0]
    e = out - T
    loss = float(np.sum(e * e) / P)
    dout = 2.0 * e / P
    offs, total = layout(W)
    g = np.zeros(total)
    A3 = c["A"][3]
    ow, ob, fo, fi = offs[3]
    g[ow:ow + fo * fi] = (dout.T @ A3).ravel()
    g[ob:ob + fo] = dout.sum(axis=0)
    dZ = r((dout @ Ws[3]) * dsilu(c["Z"][2]))
    for l in (2, 1, 0):
        ow, ob, fo, fi = offs[l]
        g[ow:ow + fo * fi] = (dZ.T @ c["A"][l]).ravel()
        g[ob:ob + fo] = dZ.sum(axis=0)
        if l > 0:
            dZ = r((dZ @ r(Ws[l])) * dsilu(c["Z"][l - 1]))
    return loss, g


def loss_only(params, X, T, W):
    out, _ = forward(params, X, W, emulate=False)
    e = out - T
    return float(np.sum(e * e) / X.shape[0])


def adam_step(params, m, v, g, step, W, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """One AdamW step; returns (params, m, v) as new arrays."""
    offs, total = layout(W)
    is_w = np.zeros(total, dtype=bool)
    for ow, ob, fo, fi in offs:
        is_w[ow:ow + fo * fi] = True
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    p = np.where(is_w, params - lr * wd * params, params)
    p = p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
