"""Synthetic viewer frames for the temporal member's checks.

A floor (z = 0) split into two leaves at x = 0, a wall (y = 3) of a third
leaf and the sky, seen by a pinhole camera that holds still, steps sideways
or orbits. Each frame gives, per pixel, the G-buffer's surface (the ray
through the pixel's centre), the record's surface (a ray jittered within
the pixel, so that it may meet another leaf at an edge) and the pixel's
light (the surface's albedo times a smooth lighting, with seeded noise).
"""
import math

import numpy as np

from oracle import temporal_ref as T

ALBEDO = {1: (0.7, 0.6, 0.5), 2: (0.2, 0.5, 0.8), 3: (0.9, 0.9, 0.0005)}


def lens(eye, target, up=(0, 0, 1), fov=0.8, width=32, height=24):
    eye = np.asarray(eye, float)
    forward = np.asarray(target, float) - eye
    forward = forward / math.sqrt(forward @ forward)
    right = np.cross(forward, up)
    right = right / math.sqrt(right @ right)
    upv = np.cross(right, forward)
    return (tuple(eye), tuple(forward), tuple(right), tuple(upv), math.tan(fov / 2), width / height)


def met(eye, direction):
    """(distance, normal facing the ray, leaf) of the nearest surface."""
    best = (math.inf, (0.0, 0.0, 0.0), T.NO_LEAF)
    if direction[2] < 0:
        t = -eye[2] / direction[2]
        x = eye[0] + t * direction[0]
        best = (t, (0.0, 0.0, 1.0), 1 if x < 0 else 2)
    if direction[1] > 0:
        t = (3 - eye[1]) / direction[1]
        if t < best[0]:
            best = (t, (0.0, -1.0, 0.0), 3)
    return best


def surface(eye, direction):
    s = np.zeros((), T.SURFACE_DTYPE)
    d, n, leaf = met(eye, direction)
    s["distance"] = d
    s["normal"] = n
    s["leaf"] = leaf
    s["albedo"] = ALBEDO.get(leaf, (0.0, 0.0, 0.0))
    return s


def frame(lens_now, lens_before, width, height, rng, moved=False, restarted=False, recounted=False, noise=0.3):
    eye = np.asarray(lens_now[0])
    cam = T.Camera(lens_now, width, height)
    record = np.zeros((height, width), T.SURFACE_DTYPE)
    gbuffer = np.zeros((height, width), T.SURFACE_DTYPE)
    parts = np.zeros((height, width, 3))
    for y in range(height):
        for x in range(width):
            h = np.asarray(cam.heading(x + 0.5, y + 0.5))
            gbuffer[y, x] = surface(eye, h / math.sqrt(h @ h))
            u, v = rng.random(2)
            hj = np.asarray(cam.heading(x + u, y + v))
            s = surface(eye, hj / math.sqrt(hj @ hj))
            record[y, x] = s
            if s["leaf"] == T.NO_LEAF:
                light = np.array([0.3, 0.4, 0.6])
            else:
                p = eye + hj / math.sqrt(hj @ hj) * float(s["distance"])
                light = np.asarray(ALBEDO[int(s["leaf"])]) * (1.0 + 0.5 * math.sin(p[0] * 1.7) * math.cos(p[1] * 1.3))
            parts[y, x] = light * (1 + noise * rng.standard_normal())
    return dict(width=width, height=height, lens=lens_now, before_lens=lens_before, moved=moved, restarted=restarted,
                recounted=recounted, parts=parts, record=record, gbuffer=gbuffer)


def sequence(seed=3, width=32, height=24, frames=8):
    """A restart, still frames, sideways steps, an orbit step and a
    recounted frame: a list of frames in the oracle's form."""
    rng = np.random.default_rng(seed)
    out = []
    target = (0.0, 1.0, 0.0)
    eye = np.array([0.0, -4.0, 2.5])
    before = lens(eye, target, width=width, height=height)
    for f in range(frames):
        moved = f >= 3
        if f >= 3 and f < 6:
            eye = eye + np.array([0.06, 0.0, 0.0])
        elif f >= 6:
            a = 0.05
            eye = np.array([eye[0] * math.cos(a) - eye[1] * math.sin(a), eye[0] * math.sin(a) + eye[1] * math.cos(a), eye[2]])
        now = lens(eye, target, width=width, height=height)
        out.append(frame(now, before, width, height, rng, moved=moved, restarted=(f == 0), recounted=(f == frames - 1)))
        before = now
    return out


def run_oracle(frames):
    """Every frame through the oracle: [(pixels, image)]."""
    state, results = None, []
    for fr in frames:
        px, img = T.temporal_frame(state, fr)
        results.append((px, img))
        state = dict(width=fr["width"], height=fr["height"], demodulated=True, pixels=px)
    return results
