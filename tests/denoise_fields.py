"""Synthetic histories for the denoiser's tests, as the reference's own
denoiser tests build them (target/tests/denoise.cpp:13-150): a 64x48 lens
looking down -z, pixels of a given light, moment, length, distance 10,
normal +z; noise from SplitMix64 of the pixel and a seed."""
import math

import numpy as np

from kodanuki_amd import DENOISE_PIXEL_DTYPE
from kodanuki_amd.scenes import mixed_bits

WIDTH, HEIGHT, EDGE = 64, 48, 32
FORWARD, RIGHT, UP, REACH = (0, 0, -1), (1, 0, 0), (0, 1, 0), 0.5


def noise(x, y, seed):
    bits = mixed_bits(x | (y << 20) | (seed << 40))
    return (bits >> 11) * 2.0 ** -52 - 1


def luminance(c):
    return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]


def depth_distance(width, height, x, y, depth):
    sideways = REACH * width / height
    across = (2 * (x + 0.5) / width - 1) * sideways
    down = (1 - 2 * (y + 0.5) / height) * REACH
    return depth * math.sqrt(1 + across * across + down * down)


def field(pixel, width=WIDTH, height=HEIGHT):
    """A history from pixel(x, y) -> dict(colour, extra, length, ...)."""
    px = np.zeros((height, width), DENOISE_PIXEL_DTYPE)
    for y in range(height):
        for x in range(width):
            d = pixel(x, y)
            c = d["colour"]
            px[y, x]["colour"] = c
            px[y, x]["moment"] = luminance(c) ** 2 + d.get("extra", 0.0)
            px[y, x]["length"] = d.get("length", 1)
            px[y, x]["distance"] = d.get("distance", 10)
            px[y, x]["normal"] = d.get("normal", (0, 0, 1))
            px[y, x]["leaf"] = d.get("leaf", 1)
            px[y, x]["albedo"] = d.get("albedo", (0, 0, 0))
    return px


def noisy_plane(width=WIDTH, height=HEIGHT):
    def pixel(x, y):
        v = 0.5 + 0.3 * noise(x, y, 1)
        return {"colour": (v, v, v)}
    return field(pixel, width, height)


def edge_field(tilted):
    def pixel(x, y):
        right = x >= EDGE
        v = (0.6 if right else 0.4) + 0.3 * noise(x, y, 2)
        return {"colour": (v, v, v), "distance": depth_distance(WIDTH, HEIGHT, x, y, 12 if right and not tilted else 10),
                "normal": (0.8, 0, 0.6) if right and tilted else (0, 0, 1)}
    return field(pixel)


def checker_albedo(x, y):
    return 0.2 if (x // 4 + y // 4) % 2 == 0 else 0.8


def checker_kept_field():
    def pixel(x, y):
        a = checker_albedo(x, y)
        v = a * (1 + 0.3 * noise(x, y, 3))
        return {"colour": (v, v, v), "albedo": (a, a, a)}
    return field(pixel)


def checker_field(demodulated):
    def pixel(x, y):
        a = checker_albedo(x, y)
        light = 1 + 0.3 * noise(x, y, 4)
        scale = 1 if demodulated else a
        v = light * scale
        return {"colour": (v, v, v), "extra": 0.0025 * scale * scale, "length": 8, "albedo": (a, a, a)}
    return field(pixel)


def converged_field():
    def pixel(x, y):
        v = 0.5 + 0.3 * math.sin(x / 5) * math.cos(y / 7)
        return {"colour": (v, v, v), "extra": 0.0025, "length": 256}
    return field(pixel)


def sky_field():
    """Sky on the top half (no leaf, infinite distance), a noisy plane below."""
    def pixel(x, y):
        v = 0.5 + 0.3 * noise(x, y, 5)
        if y < HEIGHT // 2:
            return {"colour": (v * 2, v * 2, v * 2), "distance": math.inf, "leaf": 0xFFFFFFFF, "normal": (0, 0, 0)}
        return {"colour": (v, v, v), "length": 3 if x % 7 == 0 else 6}
    return field(pixel)


def red_error(image, want):
    h, w = image.shape[:2]
    ys, xs = np.mgrid[0:h, 0:w]
    target = np.vectorize(want)(xs, ys) if callable(want) else want
    return float(np.sqrt(np.mean((image[..., 0].astype(np.float64) - target) ** 2)))


def colour_image(px):
    return px["colour"].astype(np.float32)
