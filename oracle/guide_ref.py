"""The CPU oracle of the learned guide: TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, to check the device network against it; the
product (libkodanuki_hip.so) never calls it.

A numpy restatement of the guide the library runs (the reference has no
learned guide; it joins the direction sampling of
plugins/transport/paths/sampler.h:23,357):
- the input encoding: a hash grid of 16 levels (resolutions 16 .. 128 by
  default in a geometric series, rounded to nearest; 2^19 entries a level, two
  features; levels whose (res + 1)^3 corners fit indexed densely, the
  others by the spatial hash
  x * 1 ^ y * 2654435761 ^ z * 805459861 of Instant NGP, Mueller et al.
  2022), trilinear weights; real spherical harmonics of degrees 0..3 of the
  outgoing direction; the normal; sin/cos of pi t 2^j (j < 4) for the
  wavelength t = (nm - 380) / 400; a constant one;
- four 64x64 layers, ReLU between them, no biases;
- the outputs as 8 von Mises-Fisher lobes: softmax shares, normalised
  axes, sharpness exp(clamp(c, -3, 6.2)); density
  sum share * k / (2 pi (1 - e^-2k)) e^(k (axis . w - 1)); and the
  probability a of sampling the mixture rather than the BSDF: a fixed
  share, or the share times sigmoid(output 40) when it is learned;
- the loss -weight * log p of the one-sample MIS density
  p = a q + (1 - a) pB (pB the BSDF's density of the direction), or
  -weight * log q for samples without one, and its gradient through the
  layers and the grid (Neural Parametric Mixtures, Dong et al. 2023; the
  KL gradient estimate of Neural Importance Sampling, Mueller et al. 2019).

Two modes: `bf16=True` rounds where the device rounds (the network's
inputs, each ReLU activation, the gradient into each layer; FP32 sums), for
comparison with the device within BF16 tolerances; `bf16=False` computes in
FP64 throughout, which the finite-difference checks use.
"""
import math

import numpy as np

LEVELS, FEATURES, TABLE, WIDTH, LAYERS, LOBES = 16, 2, 1 << 19, 64, 4, 8
SELECT = 5 * LOBES  # the output of the selection logit
OUTPUTS = 5 * LOBES + 1
DEPTHS, SHARE = 4, 0.3
LOG_SHARP_MIN, LOG_SHARP_MAX = -3.0, 6.2
MIN_RES, MAX_RES = 16, 128  # the device's default; 2048 and others by argument
IN_GRID, IN_SH, IN_NORMAL, IN_WAVE, IN_ONE = 0, 32, 48, 51, 59


def resolutions(max_res=MAX_RES):
    """(res per level, dense mask)."""
    growth = math.exp((math.log(max_res) - math.log(MIN_RES)) / (LEVELS - 1))
    res, dense = [], 0
    for level in range(LEVELS):
        r = int(math.floor(MIN_RES * growth ** level + 0.5))
        res.append(r)
        if (r + 1) ** 3 <= TABLE:
            dense |= 1 << level
    return res, dense


def bf16(x):
    """Round float32 values to BF16 (nearest even), as float32."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, ((u >> 16) | 0x40) << 16, r)
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).reshape(a.shape)


class Guide:
    """Parameters of a guide: grid [LEVELS, TABLE, FEATURES] and
    mlp [LAYERS, WIDTH, WIDTH] (input, output), the region."""

    def __init__(self, grid, mlp, lower, size, max_res=MAX_RES, cap=SHARE, learned=False):
        self.grid = np.asarray(grid).reshape(LEVELS, TABLE, FEATURES)
        self.mlp = np.asarray(mlp).reshape(LAYERS, WIDTH, WIDTH)
        self.lower = np.asarray(lower, np.float32)
        self.inv_size = (np.float32(1) / np.asarray(size, np.float32)).astype(np.float32)
        self.res, self.dense = resolutions(max_res)
        self.cap = cap
        self.learned = learned


def _corners(g, level, p, dt):
    """Indices [B, 8] and trilinear weights [B, 8] of a level."""
    n = g.res[level]
    s = (p * dt(n)).astype(dt)
    fl = np.clip(np.floor(s), 0, n - 1)
    f = (s - fl).astype(dt)
    c = fl.astype(np.uint64)
    idx = np.zeros((len(p), 8), np.uint64)
    wt = np.zeros((len(p), 8), dt)
    for corner in range(8):
        d = np.array([corner & 1, corner >> 1 & 1, corner >> 2 & 1], np.uint64)
        x, y, z = (c + d).T
        if g.dense >> level & 1:
            side = np.uint64(n + 1)
            idx[:, corner] = x + side * (y + side * z)
        else:
            h = (x * np.uint64(1)) ^ ((y * np.uint64(2654435761)) & np.uint64(0xFFFFFFFF)) ^ \
                ((z * np.uint64(805459861)) & np.uint64(0xFFFFFFFF))
            idx[:, corner] = h & np.uint64(TABLE - 1)
        w = np.ones(len(p), dt)
        for k in range(3):
            w = (w * np.where(d[k] == 1, f[:, k], dt(1) - f[:, k])).astype(dt)
        wt[:, corner] = w
    return idx, wt


def positions(g, x, dt=np.float32):
    t = ((np.asarray(x, dt) - g.lower.astype(dt)) * g.inv_size.astype(dt)).astype(dt)
    return np.clip(t, 0, 1).astype(dt)


def sh16(d, dt=np.float32):
    x, y, z = (np.asarray(d, dt)[:, k] for k in range(3))
    xx, yy, zz = x * x, y * y, z * z
    c = [dt(v) for v in (0.28209479177387814, 0.48860251190291987, 1.0925484305920792, 0.94617469575755997,
                         0.31539156525251999, 0.54627421529603959, 0.59004358992664352, 2.8906114426405538,
                         0.45704579946446572, 0.3731763325901154, 1.4453057213202769)]
    o = [np.full_like(x, c[0]), -c[1] * y, c[1] * z, -c[1] * x, c[2] * x * y, -c[2] * y * z, c[3] * zz - c[4],
         -c[2] * x * z, c[5] * (xx - yy), c[6] * y * (yy - 3 * xx), c[7] * x * y * z, c[8] * y * (1 - 5 * zz),
         c[9] * z * (5 * zz - 3), c[8] * x * (1 - 5 * zz), c[10] * z * (xx - yy), c[6] * x * (3 * yy - xx)]
    return np.stack(o, axis=1).astype(dt)


def inputs(g, x, out, normal, nm, dt=np.float32):
    """The network's inputs [B, 64] and each level's corners."""
    x, out, normal = (np.asarray(a, dt).reshape(-1, 3) for a in (x, out, normal))
    nm = np.asarray(nm, dt).reshape(-1)
    B = len(x)
    p = positions(g, x, dt)
    X = np.zeros((B, WIDTH), dt)
    corners = []
    for level in range(LEVELS):
        idx, wt = _corners(g, level, p, dt)
        corners.append((idx, wt))
        table = g.grid[level].astype(dt)
        f0 = np.zeros(B, dt)
        f1 = np.zeros(B, dt)
        for c in range(8):
            f0 = (f0 + wt[:, c] * table[idx[:, c], 0]).astype(dt)
            f1 = (f1 + wt[:, c] * table[idx[:, c], 1]).astype(dt)
        X[:, IN_GRID + 2 * level] = f0
        X[:, IN_GRID + 2 * level + 1] = f1
    X[:, IN_SH:IN_SH + 16] = sh16(out, dt)
    X[:, IN_NORMAL:IN_NORMAL + 3] = normal
    t = ((nm - dt(380)) * dt(1.0 / 400)).astype(dt)
    phase = (dt(np.pi) * t).astype(dt)
    for j in range(4):
        X[:, IN_WAVE + 2 * j] = np.sin(phase)
        X[:, IN_WAVE + 2 * j + 1] = np.cos(phase)
        phase = (phase * 2).astype(dt)
    X[:, IN_ONE] = 1
    return X, corners


def forward(g, X, bf=True):
    """Activations H[0..3] (the inputs of each layer) and outputs O."""
    dt = np.float32 if bf else np.float64
    W = g.mlp.astype(dt)
    if bf:
        W = bf16(W)
    H = [bf16(X) if bf else X.astype(dt)]
    for layer in range(LAYERS):
        Z = (H[-1] @ W[layer]).astype(dt)
        if layer + 1 < LAYERS:
            A = np.maximum(Z, 0)
            H.append(bf16(A) if bf else A)
        else:
            O = Z
    return H, O


def select(O, cap=SHARE, learned=False):
    """The selection probability [B]: cap, or cap * sigmoid(output 40)."""
    O = np.asarray(O)
    return cap / (1 + np.exp(-O[:, SELECT])) if learned else np.full(len(O), cap)


def decode(O):
    """shares [B, 8], axes [B, 8, 3], sharpness [B, 8]."""
    O = np.asarray(O)
    logits = O[:, 0:5 * LOBES:5]
    e = np.exp(logits - logits.max(axis=1, keepdims=True))
    share = e / e.sum(axis=1, keepdims=True)
    raw = np.stack([O[:, 5 * k + 1:5 * k + 4] for k in range(LOBES)], axis=1)
    length = np.linalg.norm(raw, axis=2, keepdims=True)
    axis = np.where(length > 1e-6, raw / np.maximum(length, 1e-30), np.array([0, 0, 1.0]))
    sharp = np.exp(np.clip(O[:, 4:5 * LOBES:5], LOG_SHARP_MIN, LOG_SHARP_MAX))
    return share, axis, sharp


def vmf_norm(k):
    return k / (2 * np.pi * (1 - np.exp(-2 * k)))


def mixture_pdf(share, axis, sharp, w):
    """q(w) for directions w [B, 3] (one per mixture) or [B, M, 3]."""
    w = np.asarray(w)
    if w.ndim == 2:
        c = np.einsum("bkj,bj->bk", axis, w)
        return np.sum(share * vmf_norm(sharp) * np.exp(sharp * (c - 1)), axis=1)
    c = np.einsum("bkj,bmj->bmk", axis, w)
    return np.sum(share[:, None] * vmf_norm(sharp)[:, None] * np.exp(sharp[:, None] * (c - 1)), axis=2)


def mixture_sample(share, axis, sharp, u_lobe, u, v):
    """Directions by lobe choice and Jakob's inversion, for one mixture."""
    cdf = np.cumsum(share)
    k = np.minimum(np.searchsorted(cdf, u_lobe, side="right"), LOBES - 1)
    s = sharp[k]
    w = 1 + np.log(u + (1 - u) * np.exp(-2 * s)) / s
    r = np.sqrt(np.maximum(0, 1 - w * w))
    phi = 2 * np.pi * v
    local = np.stack([r * np.cos(phi), r * np.sin(phi), w], axis=1)
    out = np.zeros_like(local)
    for j in range(len(local)):
        n = axis[k[j]]
        t, b = _basis(n)
        out[j] = t * local[j, 0] + b * local[j, 1] + n * local[j, 2]
    return out


def _basis(n):
    a = np.array([1.0, 0, 0]) if abs(n[0]) < 0.9 else np.array([0, 1.0, 0])
    t = np.cross(a, n)
    t /= np.linalg.norm(t)
    return t, np.cross(n, t)


def loss_gradient(O, w_in, weight, bsdf=None, cap=SHARE, learned=True):
    """Per-sample loss [B] and its gradient dL/dO [B, 64]: -weight log p of
    p = a q + (1 - a) bsdf where bsdf >= 0, -weight log q where it is
    negative or not given (zero for samples whose density underflows or
    that carry nothing)."""
    O = np.asarray(O)
    dt = O.dtype
    n = len(O)
    bsdf = np.full(n, -1.0) if bsdf is None else np.asarray(bsdf, np.float64)
    share, axis, sharp = decode(O)
    dots = np.einsum("bkj,bj->bk", axis, w_in)
    v = vmf_norm(sharp) * np.exp(sharp * (dots - 1))
    q = np.sum(share * v, axis=1)
    s = 1 / (1 + np.exp(-O[:, SELECT]))
    a = cap * s if learned else np.full(n, cap)
    mixed = bsdf >= 0
    p = np.where(mixed, a * q + (1 - a) * np.maximum(bsdf, 0), q)
    ok = (q > 1e-30) & (weight > 0) & (p > 1e-30)
    qs = np.where(ok, q, 1)
    ps = np.where(ok, p, 1)
    gamma = share * v / qs[:, None]
    w0 = np.where(ok, weight, 0).astype(dt)
    loss = np.where(ok, -w0 * np.log(ps), 0)
    G = np.zeros((n, WIDTH), dt)
    if learned:
        G[:, SELECT] = np.where(mixed & ok, -w0 * (q - np.maximum(bsdf, 0)) / ps * cap * s * (1 - s), 0)
    wgt = np.where(mixed, w0 * a * q / ps, w0).astype(dt)
    G[:, 0:5 * LOBES:5] = -wgt[:, None] * (gamma - share)
    raw = np.stack([O[:, 5 * k + 1:5 * k + 4] for k in range(LOBES)], axis=1)
    length = np.linalg.norm(raw, axis=2)
    s = np.where(length > 1e-6, -wgt[:, None] * gamma * sharp / np.maximum(length, 1e-30), 0)
    dax = s[:, :, None] * (w_in[:, None, :] - dots[:, :, None] * axis)
    for k in range(LOBES):
        G[:, 5 * k + 1:5 * k + 4] = dax[:, k]
    c = O[:, 4:5 * LOBES:5]
    inside = (c > LOG_SHARP_MIN) & (c < LOG_SHARP_MAX)
    dlogv = 1 / sharp - 2 / (np.exp(2 * sharp) - 1) + (dots - 1)
    G[:, 4:5 * LOBES:5] = np.where(inside, -wgt[:, None] * gamma * dlogv * sharp, 0)
    return loss, G


def gradients(g, samples, bf=True):
    """Loss sum and gradients (grid [LEVELS, TABLE, 2] as a dict of touched
    entries {(level, entry): [2]}, mlp [4, 64, 64]) of samples [B, 16]."""
    dt = np.float32 if bf else np.float64
    S = np.asarray(samples, np.float64 if not bf else np.float32)
    X, corners = inputs(g, S[:, 0:3], S[:, 3:6], S[:, 6:9], S[:, 9], dt)
    H, O = forward(g, X, bf)
    loss, G = loss_gradient(O, S[:, 10:13].astype(dt), S[:, 13].astype(dt), S[:, 14], g.cap, g.learned)
    W = g.mlp.astype(dt)
    if bf:
        W = bf16(W)
        G = bf16(G)
    dW = np.zeros((LAYERS, WIDTH, WIDTH), dt)
    for layer in range(LAYERS - 1, -1, -1):
        dW[layer] = (H[layer].T @ G).astype(dt)
        dH = (G @ W[layer].T).astype(dt)
        if layer > 0:
            dH = np.where(H[layer] > 0, dH, 0).astype(dt)
            G = bf16(dH) if bf else dH
        else:
            dIn = dH
    grid = {}
    for level in range(LEVELS):
        idx, wt = corners[level]
        for b in range(len(S)):
            for c in range(8):
                key = (level, int(idx[b, c]))
                add = wt[b, c] * dIn[b, IN_GRID + 2 * level:IN_GRID + 2 * level + 2]
                grid[key] = grid.get(key, 0) + add
    return float(np.sum(loss, dtype=np.float64)), grid, dW


def loss_of(g, samples):
    """FP64 loss sum of samples (for finite differences)."""
    S = np.asarray(samples, np.float64)
    X, _ = inputs(g, S[:, 0:3], S[:, 3:6], S[:, 6:9], S[:, 9], np.float64)
    _, O = forward(g, X, bf=False)
    loss, _ = loss_gradient(O, S[:, 10:13], S[:, 13], S[:, 14], g.cap, g.learned)
    return float(np.sum(loss))
