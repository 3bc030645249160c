"""Test infrastructure: the oracle of a pixel's estimate and of the image
outputs, restated in numpy from the reference. Only tests/ load it.

- fold / fold_pixel: foldedSample and closedBatch
  (source/core/stopping.cpp:36-67) for samples drawn alone (split 1: a
  sample's shared part is its whole colour, core/stopping.h:21-32), each
  operation in the reference's order (a batch's mean is its sum times the
  reciprocal of the batch, its shared part the shared sum over the batch);
  channelMean is (r + g + b) / 3 (source/math/colour.cpp:53-55).
- displayed / encode: the grade (source/shared/image.cpp:25-53): value
  times exposure, the filmic curve (saturating at 16) or clamped, clamped
  to [0, 1], to the power 1/2.2, times 255, rounded half away from zero.
- ppm_bytes / pfm_bytes: the files writePpm and writePfm write
  (source/shared/image.cpp:55-96).
"""
import math

import numpy as np

FIELDS = ("total", "taken", "batches", "batch", "means", "squares", "shared", "shared_means", "shared_squares", "own_squares")


def empty():
    return {"total": [0.0, 0.0, 0.0], "taken": 0, "batches": 0, "batch": [0.0, 0.0, 0.0], "means": [0.0, 0.0, 0.0],
            "squares": [0.0, 0.0, 0.0], "shared": 0.0, "shared_means": 0.0, "shared_squares": 0.0, "own_squares": 0.0}


def channel_mean(c):
    return (c[0] + c[1] + c[2]) / 3


def fold(e, colour, batch):
    """One sample drawn alone into estimate e (a dict of Python floats)."""
    c = [float(x) for x in colour]
    e = dict(e)
    e["total"] = [e["total"][k] + c[k] for k in range(3)]
    e["taken"] = e["taken"] + 1
    e["batch"] = [e["batch"][k] + c[k] for k in range(3)]
    e["shared"] = e["shared"] + channel_mean(c)
    if batch > 0 and e["taken"] % batch == 0:
        scale = 1.0 / batch
        mean = [e["batch"][k] * scale for k in range(3)]
        shared = e["shared"] / batch
        own = channel_mean(mean) - shared
        e["means"] = [e["means"][k] + mean[k] for k in range(3)]
        e["squares"] = [e["squares"][k] + mean[k] * mean[k] for k in range(3)]
        e["batch"] = [0.0, 0.0, 0.0]
        e["batches"] = e["batches"] + 1
        e["shared_means"] = e["shared_means"] + shared
        e["shared_squares"] = e["shared_squares"] + shared * shared
        e["own_squares"] = e["own_squares"] + own * own
        e["shared"] = 0.0
    return e


def fold_pixel(e, colours, batch):
    for c in colours:
        e = fold(e, c, batch)
    return e


def as_record(e, dtype):
    r = np.zeros((), dtype)
    for f in FIELDS:
        r[f] = e[f]
    return r


def from_record(r):
    return {f: (r[f].tolist() if np.ndim(r[f]) else r[f].item()) for f in FIELDS}


FILMIC_CEILING = 2.51 / 2.43
FILMIC_SATURATED = 16.0


def filmic(v):
    if v >= FILMIC_SATURATED:
        return FILMIC_CEILING
    return v * (2.51 * v + 0.03) / (v * (2.43 * v + 0.59) + 0.14)


def displayed(value, exposure=1.0, curve=0):
    if not value > 0:
        return 0.0
    exposed = value * exposure
    mapped = filmic(exposed) if curve == 0 else exposed
    share = 0.0 if not mapped > 0 else min(mapped, 1.0)
    return math.pow(share, 1 / 2.2)


def round_half_away(x):
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def encode(rgb, exposure=1.0, curve=0):
    """Bytes of a float32 image (the value widened to double first)."""
    flat = np.asarray(rgb, np.float32).ravel()
    out = np.array([round_half_away(displayed(float(v), float(np.float32(exposure)), curve) * 255.0) for v in flat], np.uint8)
    return out.reshape(np.shape(rgb))


def ppm_bytes(rgb, exposure=1.0, curve=0):
    h, w, _ = np.shape(rgb)
    return f"P6\n{w} {h}\n255\n".encode() + encode(rgb, exposure, curve).tobytes()


def pfm_bytes(rgb):
    img = np.asarray(rgb, np.float32)
    h, w, _ = img.shape
    return f"PF\n{w} {h}\n-1\n".encode() + b"".join(img[y].astype("<f4").tobytes() for y in range(h - 1, -1, -1))
