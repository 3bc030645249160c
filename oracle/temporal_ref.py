"""Temporal reprojection in plain Python: TEST INFRASTRUCTURE ONLY.

Only tests/ imports this module, to check the device's temporal step; the
product never does.

Restates the reference's temporal member (plugins/reconstruction/temporal/
temporal.cpp) for the frames a viewer draws with a G-buffer at factor 1:
every pixel drawn each frame through its centre, the history demodulated
(temporalFrame, temporal.h:237, with demodulatedFrame true). Per pixel:
the 3x3 neighbourhood of the frame's demodulated sample lights (mean and
deviation: rowNeighbourhoods, temporal.cpp:935-955); the history, none
after a restart or a resize, the pixel's own on a still camera, else the
G-buffer's point reprojected into the camera before and read bilinearly
from the pixels on its plane (surfacedHistory, temporal.cpp:584-603); the
blend with the frame's sample where its first surface is the G-buffer's
leaf, the history clamped to the neighbourhood (heldBlend,
temporal.cpp:846-866), else the fresh fill (jointFilled,
temporal.cpp:754-765); the shown light the lighting times the G-buffer's
albedo. Python floats are IEEE doubles and every expression keeps the
reference's order of operations, so results are meant to be bit-equal.

Pinned by tests/test_temporal.py against the reference's own temporalFrame
(built from its temporal.cpp into oracle/_ref) on seeded frames, against
fixtures written from it, and by the reference test's behaviours
(target/tests/reconstruction.cpp).
"""
import math

import numpy as np

NO_LEAF = 0xFFFFFFFF
MOVING_CAP, STILL_CAP, YOUNG = 32.0, 256.0, 8.0
MOVING_SPREAD, STILL_SPREAD = 1.25, 6.0
DEPTH_SHARE, NORMAL_COSINE, ALBEDO_FLOOR = 0.03, 0.9, 1e-3

SURFACE_DTYPE = np.dtype([("distance", "<f4"), ("normal", "<f4", 3), ("leaf", "<u4"), ("albedo", "<f4", 3)])
PIXEL_DTYPE = np.dtype([("colour", "<f8", 3), ("moment", "<f8"), ("length", "<f4"), ("distance", "<f4"), ("normal", "<f8", 3),
                        ("leaf", "<u4"), ("pad", "<u4"), ("albedo", "<f8", 3)])


def f32(x):
    return float(np.float32(x))


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def mul(a, s):
    return (a[0] * s, a[1] * s, a[2] * s)


def cmul(a, b):
    return (a[0] * b[0], a[1] * b[1], a[2] * b[2])


def length(a):
    return math.sqrt(dot(a, a))


def luminance(c):
    return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]


class Camera:
    """TemporalCamera of a lens at the grid's size (temporal.cpp:307-322)."""

    def __init__(self, lens, width, height):
        eye, forward, right, up, reach, aspect = lens
        sideways = reach * aspect
        self.eye, self.forward = tuple(eye), tuple(forward)
        self.across = mul(right, 2 * sideways / float(width))
        self.down = mul(up, -2 * reach / float(height))
        self.corner = add(sub(forward, mul(right, sideways)), mul(up, reach))
        self.across_inverse = mul(self.across, 1 / dot(self.across, self.across))
        self.down_inverse = mul(self.down, 1 / dot(self.down, self.down))
        self.width, self.height = float(width), float(height)

    def heading(self, x, y):
        return add(add(self.corner, mul(self.across, x)), mul(self.down, y))


def divisor(albedo):
    return tuple(1.0 if a < ALBEDO_FLOOR else a for a in albedo)


def divided(light, d):
    return (light[0] / d[0], light[1] / d[1], light[2] / d[2])


def sky_surface(s):
    return int(s["leaf"]) == NO_LEAF or not math.isfinite(float(s["distance"]))


def shown_leaf(s):
    return NO_LEAF if sky_surface(s) else int(s["leaf"])


def plane_through(s, eye, heading):
    if sky_surface(s):
        return (NO_LEAF, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.0)
    d = float(s["distance"])
    point = add(eye, mul(heading, d / length(heading)))
    return (int(s["leaf"]), point, tuple(float(v) for v in s["normal"]), DEPTH_SHARE * d)


def same_leaf(a, b, recounted):
    return ((a == NO_LEAF) == (b == NO_LEAF)) if recounted else a == b


def on_plane(plane, other, recounted):
    if not same_leaf(other[0], plane[0], recounted):
        return False
    if plane[0] == NO_LEAF:
        return True
    if dot(other[2], plane[2]) < NORMAL_COSINE:
        return False
    return abs(dot(sub(other[1], plane[1]), plane[2])) <= plane[3]


def place_before(cam, along):
    ahead = dot(along, cam.forward)
    if not ahead > 0:
        return None
    scale = 1 / ahead
    x = dot(along, cam.across_inverse) * scale + 0.5 * cam.width
    y = dot(along, cam.down_inverse) * scale + 0.5 * cam.height
    if not (x >= 0 and x < cam.width and y >= 0 and y < cam.height):
        return None
    return (x, y, length(along))


def temporal_frame(before, frame):
    """One frame. before: None or dict(width, height, demodulated, pixels
    [h, w] PIXEL_DTYPE); frame: dict(width, height, lens, before_lens,
    moved, restarted, recounted, parts [h, w, 3] f64, record and gbuffer
    [h, w] SURFACE_DTYPE), a lens being (eye, forward, right, up, reach,
    aspect). Returns (pixels [h, w] PIXEL_DTYPE, image [h, w, 3] f32)."""
    W, H = frame["width"], frame["height"]
    moved, recounted = bool(frame["moved"]), bool(frame["recounted"])
    rec, gb, parts = frame["record"], frame["gbuffer"], frame["parts"]
    kept = (before is not None and not frame["restarted"] and before["width"] == W and before["height"] == H and W > 0 and
            bool(before["demodulated"]))
    cam = Camera(frame["lens"], W, H)
    prev = Camera(frame["before_lens"], W, H)
    cx = cy = 0.5

    def sample_light(x, y):
        p = parts[y, x]
        return divided((float(p[0]), float(p[1]), float(p[2])), divisor([float(a) for a in rec[y, x]["albedo"]]))

    lights = [[sample_light(x, y) for x in range(W)] for y in range(H)]

    def clamped(x, y):
        return min(max(x, 0), W - 1), min(max(y, 0), H - 1)

    def sample_sum(x, y):
        c = lights[y][x]
        return (c, cmul(c, c))

    def joined(a, b):
        return (add(a[0], b[0]), add(a[1], b[1]))

    # the 3x3 neighbourhood's mean and deviation of every pixel
    near = {}
    for y in range(H):
        columns = []
        for x in range(W):
            mid = joined(sample_sum(*clamped(x, y)), sample_sum(*clamped(x, y - 1)))
            columns.append(joined(mid, sample_sum(*clamped(x, y + 1))))
        for x in range(W):
            s = joined(joined(columns[x], columns[max(x - 1, 0)]), columns[min(x + 1, W - 1)])
            mean = mul(s[0], 1.0 / 9)
            dev = []
            for c in range(3):
                m = s[0][c] / 9
                dev.append(math.sqrt(max(s[1][c] / 9 - m * m, 0.0)))
            near[(x, y)] = (mean, tuple(dev))

    bp = before["pixels"] if kept else None

    def last_plane(x, y):
        last = bp[y, x]
        heading = prev.heading(x + cx, y + cy)
        if int(last["leaf"]) == NO_LEAF or not math.isfinite(float(last["distance"])):
            return (NO_LEAF, (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.0)
        point = add(prev.eye, mul(heading, float(last["distance"]) / length(heading)))
        return (int(last["leaf"]), point, tuple(float(v) for v in last["normal"]), 0.0)

    def fetched(place, plane):
        ax, ay = place[0] - cx, place[1] - cy
        lx, ly = math.floor(ax), math.floor(ay)
        tx, ty = ax - lx, ay - ly
        colour, albedo, moment, length_, weight = (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 0.0, 0.0, 0.0
        for j in range(2):
            for i in range(2):
                x, y = lx + i, ly + j
                w = ((1 - tx) if i == 0 else tx) * ((1 - ty) if j == 0 else ty)
                if not (w > 0 and 0 <= x < W and 0 <= y < H):
                    continue
                last = bp[y, x]
                if not (float(last["length"]) > 0 and on_plane(plane, last_plane(x, y), recounted)):
                    continue
                colour = add(colour, mul(tuple(float(v) for v in last["colour"]), w))
                albedo = add(albedo, mul(tuple(float(v) for v in last["albedo"]), w))
                moment += float(last["moment"]) * w
                length_ = f32(length_ + f32(float(last["length"]) * w))
                weight += w
        if not weight > 0:
            return None
        return (mul(colour, 1 / weight), moment / weight, f32(length_ / weight))

    def history_for(x, y):
        if not kept:
            return None
        if not moved:
            last = bp[y, x]
            if not float(last["length"]) > 0:
                return None
            return (tuple(float(v) for v in last["colour"]), float(last["moment"]), float(last["length"]))
        heading = cam.heading(x + cx, y + cy)
        plane = plane_through(gb[y, x], cam.eye, heading)
        along = heading if plane[0] == NO_LEAF else sub(plane[1], prev.eye)
        place = place_before(prev, along)
        if place is None:
            return None
        return fetched(place, plane)

    def joint_filled(x, y):
        heading = cam.heading(x + cx, y + cy)
        plane = plane_through(gb[y, x], cam.eye, heading)

        def share(sx, sy, w):
            if not w > 0:
                return None
            other = plane_through(rec[sy, sx], cam.eye, cam.heading(sx + cx, sy + cy))
            if not on_plane(plane, other, False):
                return None
            return mul(lights[sy][sx], w)

        corner = share(x, y, 1.0)
        if corner is not None:
            return mul(corner, 1 / 1.0)
        total, weight = (0.0, 0.0, 0.0), 0.0
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                got = share(*clamped(x + dx, y + dy), 1.0)
                if got is not None:
                    total = add(total, got)
                    weight += 1.0
        if weight > 0:
            return mul(total, 1 / weight)
        # filledColour: the bilinear blend of the four drawn samples around
        # the pixel's centre, which at factor 1 sits on the pixel itself
        xr, yb = min(x + 1, W - 1), min(y + 1, H - 1)
        top = add(mul(lights[y][x], 1 - 0.0), mul(lights[y][xr], 0.0))
        bottom = add(mul(lights[yb][x], 1 - 0.0), mul(lights[yb][xr], 0.0))
        return add(mul(top, 1 - 0.0), mul(bottom, 0.0))

    cap = MOVING_CAP if moved else STILL_CAP
    spread = MOVING_SPREAD if moved else STILL_SPREAD
    out = np.zeros((H, W), PIXEL_DTYPE)
    image = np.zeros((H, W, 3), np.float32)
    for y in range(H):
        for x in range(W):
            g = gb[y, x]
            aimed = shown_leaf(rec[y, x]) == shown_leaf(g)
            hist = history_for(x, y)
            sample = lights[y][x]
            if hist is not None:
                colour, moment, hlen = hist
                length_ = min(hlen, cap)
                if not moved and not hlen < YOUNG:
                    held = colour
                else:
                    mean, dev = near[(x, y)]
                    held = tuple(min(max(colour[c], mean[c] - spread * dev[c]), mean[c] + spread * dev[c]) for c in range(3))
                if aimed:
                    lum = luminance(sample)
                    alpha = max(1.0 / (min(hlen, cap) + 1), 1.0 / cap)
                    colour = add(mul(held, 1 - alpha), mul(sample, alpha))
                    moment = moment * (1 - alpha) + lum * lum * alpha
                    length_ = min(f32(length_ + 1), cap)
                else:
                    colour = held
            else:
                colour = sample if aimed else joint_filled(x, y)
                lum = luminance(colour)
                moment = lum * lum
                length_ = 1.0 if aimed else 0.0
            albedo = tuple(float(a) for a in g["albedo"])
            px = out[y, x]
            px["colour"] = colour
            px["moment"] = moment
            px["length"] = length_
            px["distance"] = g["distance"]
            px["normal"] = [float(v) for v in g["normal"]]
            px["leaf"] = g["leaf"]
            px["albedo"] = albedo
            shown = cmul(colour, divisor(albedo))
            image[y, x] = shown
    return out, image
