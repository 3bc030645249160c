"""The CPU oracle of the denoiser: TEST INFRASTRUCTURE ONLY.

Only tests/ import this module, to check the device filter against it; the
product (libkodanuki_hip.so) never calls it.

A numpy restatement of the reference's à-trous filter,
denoisedFrame (source/plugins/reconstruction/atrous/atrous.cpp:18-500,
documented at atrous.h:430-458), pixel arrays at a time, each operation
in the reference's precision (FP32 for the filter, FP64 for the guides'
points and the history's variance) and order, so that its results are the
reference's bits:
- preparedDenoise (atrous.cpp:60-160): each pixel's point along its ray
  through the pinhole (the lens's heading to the pixel's centre), normal,
  plane offset, inverse reach 1 / (0.02 distance), sky flag; its light
  divided by the albedo (temporal.cpp:208-214: channels under 1e-3 by 1)
  and the variance of its history (length >= 4), else -1;
- estimatedChunk (atrous.cpp:333-380): a young pixel's variance from the
  7x7 pixels around it, weighed by plane and normal;
- filteredChunk (atrous.cpp:390-420) 4 times at steps 1, 2, 4, 8: the
  deviation through a 3x3 gaussian, the 5x5 B3 spline taps weighed by the
  normals' cosine to the 128th and quickExp(-plane - apart) (atrous.cpp:
  27-35, 240-296);
- remodulatedImage (atrous.cpp:446-470): times the albedo, faded into the
  held light by 1 / (1 + length / 8).
"""
import numpy as np

F = np.float32
PASSES, VARIANCE_RADIUS, YOUNG = 4, 3, F(4)
SETTLED, PLANE_SHARE = 8.0, 0.02
SPREAD, LUM_FLOOR, LUM_SHARE, EXP_FLOOR, FACING_FLOOR = F(4), F(1e-4), F(1e-3), F(-30), F(0.5)
ALBEDO_FLOOR = 1e-3
SPLINE = [F(1.0 / 16), F(1.0 / 4), F(3.0 / 8), F(1.0 / 4), F(1.0 / 16)]
GAUSS = [F(0.25), F(0.5), F(0.25)]


def quick_exp(x):
    """quickExp (atrous.cpp:27-35) on float32 arrays."""
    x = np.asarray(x, F)
    t = (np.maximum(x, EXP_FLOOR) * F(1.44269504)).astype(F)
    whole = np.floor(t).astype(F)
    f = (t - whole).astype(F)
    power = (F(1) + f * (F(0.66023) + f * F(0.33977))).astype(F)
    bits = power.view(np.int32).astype(np.int64) + (whole.astype(np.int64) << 23)
    value = (bits & 0xFFFFFFFF).astype(np.uint32).view(F)
    return np.where(x > EXP_FLOOR, value, F(0)).astype(F)


def rgb_lum(r, g, b):
    return (F(0.2126) * r + F(0.7152) * g + F(0.0722) * b).astype(F)


def lens(forward, right, up, reach, width, height, centre=(0.5, 0.5)):
    forward, right, up = (np.asarray(v, np.float64) for v in (forward, right, up))
    aspect = width / height
    sideways = reach * aspect
    return dict(corner=forward - right * sideways + up * reach, across=right * (2 * sideways / width),
                down=up * (-2 * reach / height), centre=centre, width=width, height=height)


def prepare(pixels, L, demodulated=False):
    """Guides and light of every pixel (dict of [H, W] float32 arrays)."""
    H, W = pixels.shape
    ys, xs = np.mgrid[0:H, 0:W]
    px = pixels
    dist = px["distance"].astype(np.float64)
    sky = (px["leaf"] == 0xFFFFFFFF) | ~np.isfinite(px["distance"]) | ~(px["distance"] > 0)
    h = [L["corner"][k] + L["across"][k] * (xs + L["centre"][0]) + L["down"][k] * (ys + L["centre"][1]) for k in range(3)]
    length = np.sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2])
    with np.errstate(all="ignore"):
        scale = dist / length
        p = [h[k] * scale for k in range(3)]
        n = [px["normal"][..., k] for k in range(3)]
        G = {"x": p[0].astype(F), "y": p[1].astype(F), "z": p[2].astype(F), "nx": n[0].astype(F), "ny": n[1].astype(F),
             "nz": n[2].astype(F), "offset": (p[0] * n[0] + p[1] * n[1] + p[2] * n[2]).astype(F),
             "reach": (1 / (PLANE_SHARE * dist)).astype(F), "sky": np.zeros((H, W), F)}
    for key in G:
        G[key] = np.where(sky, F(0), G[key]).astype(F)
    G["sky"] = np.where(sky, F(1), F(0)).astype(F)
    albedo = px["albedo"]
    div = np.ones((H, W, 3)) if demodulated else np.where(albedo < ALBEDO_FLOOR, 1.0, albedo)
    col = px["colour"]
    light = {"r": (col[..., 0] / div[..., 0]).astype(F), "g": (col[..., 1] / div[..., 1]).astype(F),
             "b": (col[..., 2] / div[..., 2]).astype(F)}
    light["lum"] = rgb_lum(light["r"], light["g"], light["b"])
    lum = 0.2126 * col[..., 0] + 0.7152 * col[..., 1] + 0.0722 * col[..., 2]
    s = 0.2126 * div[..., 0] + 0.7152 * div[..., 1] + 0.0722 * div[..., 2]
    spread = np.maximum(px["moment"] - lum * lum, 0.0)
    with np.errstate(all="ignore"):
        var = (spread / (px["length"].astype(np.float64) * s * s)).astype(F)
    light["var"] = np.where(px["length"] >= YOUNG, var, F(-1)).astype(F)
    return G, light


def _shift(a, dx, dy):
    """a[y + dy, x + dx] with a validity mask (zero outside)."""
    H, W = a.shape
    out = np.zeros_like(a)
    valid = np.zeros((H, W), bool)
    y0, y1 = max(0, -dy), min(H, H - dy)
    x0, x1 = max(0, -dx), min(W, W - dx)
    if y0 < y1 and x0 < x1:
        out[y0:y1, x0:x1] = a[y0 + dy:y1 + dy, x0 + dx:x1 + dx]
        valid[y0:y1, x0:x1] = True
    return out, valid


def _tap_weight(G, light, dx, dy, kernel, inv_plane, inv_dev):
    """The weight of the tap (dx, dy) for every pixel, and the tap's light."""
    q = {k: _shift(v, dx, dy)[0] for k, v in G.items()}
    ql = {k: _shift(v, dx, dy)[0] for k, v in light.items()}
    valid = _shift(G["sky"], dx, dy)[1]
    facing = (G["nx"] * q["nx"] + G["ny"] * q["ny"] + G["nz"] * q["nz"] + G["sky"] * q["sky"]).astype(F)
    power = np.where(facing > FACING_FLOOR, facing, F(0)).astype(F)
    for _ in range(7):
        power = (power * power).astype(F)
    along = (q["x"] * G["nx"] + q["y"] * G["ny"] + q["z"] * G["nz"]).astype(F)
    plane = (np.abs(along - G["offset"]) * inv_plane).astype(F)
    with np.errstate(all="ignore"):
        apart = (np.abs(ql["lum"] - light["lum"]) * inv_dev).astype(F)
        w = (kernel * power * quick_exp((-plane - apart).astype(F))).astype(F)
    return np.where(valid, w, F(0)).astype(F), ql, valid


def estimate(G, light):
    """Young pixels' variance from the 7x7 pixels around them."""
    sums = {k: np.zeros_like(light["r"]) for k in ("r", "g", "b", "sq", "w")}
    inv_plane = (G["reach"] * F(1)).astype(F)
    zero = np.zeros_like(light["r"])
    for j in range(-VARIANCE_RADIUS, VARIANCE_RADIUS + 1):
        for i in range(-VARIANCE_RADIUS, VARIANCE_RADIUS + 1):
            w, ql, valid = _tap_weight(G, light, i, j, F(1), inv_plane, zero)
            for c in ("r", "g", "b"):
                sums[c] = np.where(valid, sums[c] + w * ql[c], sums[c]).astype(F)
            sums["w"] = np.where(valid, sums["w"] + w, sums["w"]).astype(F)
            sums["sq"] = np.where(valid, sums["sq"] + w * ql["lum"] * ql["lum"], sums["sq"]).astype(F)
    with np.errstate(all="ignore"):
        mean = (rgb_lum(sums["r"], sums["g"], sums["b"]) / sums["w"]).astype(F)
        est = np.maximum((sums["sq"] / sums["w"] - mean * mean).astype(F), F(0)).astype(F)
    est = np.where(sums["w"] > 0, est, F(0)).astype(F)
    out = dict(light)
    out["var"] = np.where(light["var"] < 0, est, light["var"]).astype(F)
    return out


def filter_pass(G, light, step):
    """One pass of the 5x5 B3 spline at a step."""
    var = np.zeros_like(light["r"])
    wsum = np.zeros_like(light["r"])
    for j in range(3):
        for i in range(3):
            v, valid = _shift(light["var"], i - 1, j - 1)
            k = (GAUSS[i] * GAUSS[j]).astype(F) if hasattr(GAUSS[i] * GAUSS[j], "astype") else F(GAUSS[i] * GAUSS[j])
            var = np.where(valid, var + k * v, var).astype(F)
            wsum = np.where(valid, wsum + k, wsum).astype(F)
    with np.errstate(all="ignore"):
        dev = np.sqrt((var / wsum).astype(F)).astype(F)
        inv_dev = (F(1) / (SPREAD * dev + LUM_FLOOR + LUM_SHARE * light["lum"])).astype(F)
    inv_plane = (G["reach"] * (F(1) / F(step))).astype(F)
    sums = {k: np.zeros_like(light["r"]) for k in ("r", "g", "b", "var", "w")}
    for j in range(-2, 3):
        for i in range(-2, 3):
            kernel = F(SPLINE[i + 2] * SPLINE[j + 2])
            w, ql, valid = _tap_weight(G, light, i * step, j * step, kernel, inv_plane, inv_dev)
            for c in ("r", "g", "b"):
                sums[c] = np.where(valid, sums[c] + w * ql[c], sums[c]).astype(F)
            sums["var"] = np.where(valid, sums["var"] + w * w * ql["var"], sums["var"]).astype(F)
            sums["w"] = np.where(valid, sums["w"] + w, sums["w"]).astype(F)
    ok = sums["w"] > 0
    with np.errstate(all="ignore"):
        scale = (F(1) / sums["w"]).astype(F)
        r, g, b = ((sums[c] * scale).astype(F) for c in ("r", "g", "b"))
        out = {"r": r, "g": g, "b": b, "lum": rgb_lum(r, g, b), "var": (sums["var"] * scale * scale).astype(F)}
    return {k: np.where(ok, out[k], light[k]).astype(F) for k in out}


def remodulate(light, pixels, demodulated=False):
    px = pixels
    share = 1 / (1 + px["length"].astype(np.float64) / SETTLED)
    albedo = px["albedo"]
    div = np.where(albedo < ALBEDO_FLOOR, 1.0, albedo)
    held = px["colour"] * div if demodulated else px["colour"]
    filt = np.stack([light["r"], light["g"], light["b"]], axis=-1).astype(np.float64)
    return (filt * div * share[..., None] + held * (1 - share[..., None])).astype(F)


def denoise(pixels, L, demodulated=False):
    """The denoised image [H, W, 3] of a history [H, W] (DENOISE_PIXEL_DTYPE)."""
    G, light = prepare(pixels, L, demodulated)
    light = estimate(G, light)
    for k in range(PASSES):
        light = filter_pass(G, light, 1 << k)
    return remodulate(light, pixels, demodulated)
