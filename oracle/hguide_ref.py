"""The CPU oracle of the plain (highlight) guide's lumps: TEST
INFRASTRUCTURE ONLY.

Only tests/ import this module, to check the device's lump arithmetic
against it; the product (libkodanuki_hip.so) never calls it.

A numpy restatement of the reference's lump arithmetic
(source/plugins/sampler/highlight_guide), in FP64 as the reference
computes it, from the FP32 fields the reference keeps:
- unpack: unpackedLump (guide.cpp:52-72): the axis normalised, across
  made orthogonal to it and normalised, up = axis x across, precision the
  inverse of the covariance L L^T of the factor L = [[f0, 0], [f1, f2]]
  (FP32), height 1 / (2 pi f0 f2) (FP32); unpackGuide (guide.cpp:81-92):
  the weights over their total;
- centre: shiftedCentre (guide.cpp:131-139): the mean moved by where the
  lump's point lies, seen from the vertex, in the lump's plane;
- lump_at: lumpAt (guide.cpp:210-221): the Gaussian over the plane one
  unit along the axis as a density of directions (the 1 / along^3 of the
  map from the plane to the sphere), cut at 60 and divided by the mass
  kept, 1 - e^-30;
- density: lumpDensity (guide.cpp:227-237), zero unless both directions
  leave the normal's side;
- draw: pickedLump and drawnFromLump (guide.cpp:239-252): Box-Muller
  within the cut, then through the factor.
"""
import numpy as np

LUMPS, LUMP_FLOATS, HEAD = 12, 15, 4
FLOATS = HEAD + LUMPS * LUMP_FLOATS
CUTOFF = 60.0
KEPT = -np.expm1(-30.0)
F = np.float32


def _normalized(v):
    v = np.asarray(v, np.float64)
    return v / np.sqrt(v @ v)


def pack(lumps, material=0, share=0.5):
    """A packed guide from dicts of axis, across, point, mean, factor, weight."""
    g = np.zeros(FLOATS, F)
    g[0], g[1], g[2] = len(lumps), material, share
    for k, L in enumerate(lumps):
        g[HEAD + k * LUMP_FLOATS:HEAD + (k + 1) * LUMP_FLOATS] = np.concatenate(
            [L["axis"], L["across"], L["point"], L.get("mean", (0, 0)), L["factor"], [L["weight"]]]).astype(F)
    return g


def unpack(g):
    """The guide's lumps as the reference unpacks them (list of dicts)."""
    count = int(g[0])
    out = []
    total = 0.0
    for k in range(count):
        p = np.asarray(g[HEAD + k * LUMP_FLOATS:HEAD + (k + 1) * LUMP_FLOATS], F)
        axis = _normalized(p[0:3].astype(np.float64))
        raw = p[3:6].astype(np.float64)
        across = _normalized(raw - axis * (raw @ axis))
        f0, f1, f2 = (float(x) for x in p[11:14])
        c00, c01, c11 = f0 * f0, f0 * f1, f1 * f1 + f2 * f2
        root = f0 * f2
        det = root * root
        out.append(dict(axis=axis, across=across, up=np.cross(axis, across), point=p[6:9].astype(np.float64),
                        mean=p[9:11].astype(np.float64), precision=np.array([c11 / det, -c01 / det, c00 / det], F).astype(np.float64),
                        factor=p[11:14].astype(np.float64), height=float(F(1 / (2 * np.pi * root))), weight=float(p[14])))
        total += float(p[14])
    for L in out:
        L["weight"] = float(F(L["weight"] / total))
    return out


def centre(L, place):
    to = _normalized(L["point"] - np.asarray(place, np.float64))
    along = to @ L["axis"]
    if not along > 1e-3:
        return L["mean"].copy()
    return np.array([F(L["mean"][0] + (to @ L["across"]) / along), F(L["mean"][1] + (to @ L["up"]) / along)], np.float64)


def lump_at(L, c, d):
    """Density of directions d [n, 3] of one lump whose centre is c."""
    d = np.asarray(d, np.float64).reshape(-1, 3)
    along = d @ L["axis"]
    out = np.zeros(len(d))
    ok = along > 1e-9
    u = (d[ok] @ L["across"]) / along[ok] - c[0]
    v = (d[ok] @ L["up"]) / along[ok] - c[1]
    P = L["precision"]
    spread = P[0] * u * u + 2 * P[1] * u * v + P[2] * v * v
    val = L["height"] / KEPT * np.exp(-0.5 * spread) / along[ok] ** 3
    out[ok] = np.where(spread > CUTOFF, 0.0, val)
    return out


def density(g, place, normal, out, dirs):
    """The lumps' density of directions dirs [n, 3] at a vertex."""
    d = np.asarray(dirs, np.float64).reshape(-1, 3)
    n = np.asarray(normal, np.float64)
    lumps = unpack(g)
    total = np.zeros(len(d))
    for L in lumps:
        total += L["weight"] * lump_at(L, centre(L, place), d)
    above = (d @ n > 0) & (np.asarray(out, np.float64) @ n > 0)
    return np.where(above, total, 0.0)


def draw(g, place, numbers):
    """Directions [n, 3] drawn by (pick, u, v) rows of numbers."""
    lumps = unpack(g)
    q = np.asarray(numbers, np.float64).reshape(-1, 3)
    # pickedLump: the first lump whose running sum of weights passes pick
    below = np.cumsum([L["weight"] for L in lumps])
    index = np.minimum(np.searchsorted(below, q[:, 0], side="right"), len(lumps) - 1)
    radius = np.sqrt(-2 * np.log(np.maximum(1 - q[:, 1] * KEPT, 1e-300)))
    angle = 2 * np.pi * q[:, 2]
    a, b = radius * np.cos(angle), radius * np.sin(angle)
    res = np.zeros((len(q), 3))
    for k, L in enumerate(lumps):
        m = index == k
        c = centre(L, place)
        f = L["factor"]
        uu = c[0] + f[0] * a[m]
        vv = c[1] + f[1] * a[m] + f[2] * b[m]
        d = L["axis"][None, :] + L["across"][None, :] * uu[:, None] + L["up"][None, :] * vv[:, None]
        res[m] = d / np.sqrt(np.sum(d * d, axis=1))[:, None]
    return res
