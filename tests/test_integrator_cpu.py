"""The oracle's spectral pieces against closed forms (CPU only):
a flat spectrum is white (target/tests/spectral.cpp:11-29), the wavelength
density integrates to one over the drawn wavelengths, and RGB survives
the round trip through the smooth-step bases (spectral.cpp:31-55); the
counter-based numbers are the reference's SplitMix64 (math/random.cpp)."""
import ctypes as C

import numpy as np

from tests import oracle_py as O


def _lib():
    L = O.lib()
    L.or_sampled_wavelengths.argtypes = [C.c_double, C.c_void_p, C.c_void_p]
    L.or_colour_of.argtypes = [C.c_void_p] * 4
    L.or_spectrum_of.argtypes = [C.c_void_p] * 3
    return L


def test_flat_spectrum_is_white():
    L = _lib()
    nm, d, rgb, acc = np.zeros(4), np.zeros(4), np.zeros(3), np.zeros(3)
    n = 4096
    for i in range(n):
        L.or_sampled_wavelengths((i + 0.5) / n, O.ptr(nm), O.ptr(d))
        s = np.ones(4)
        L.or_colour_of(O.ptr(nm), O.ptr(d), O.ptr(s), O.ptr(rgb))
        acc += rgb
    assert np.allclose(acc / n, 1.0, atol=2e-3)


def test_wavelength_density_integrates_to_one():
    L = _lib()
    nm, d = np.zeros(4), np.zeros(4)
    n = 8192
    total = 0.0
    for i in range(n):
        L.or_sampled_wavelengths((i + 0.5) / n, O.ptr(nm), O.ptr(d))
        assert np.all((nm >= 380) & (nm < 700))
        total += np.mean(1.0 / d)  # E[1/p] over the draw is the range length
    assert abs(total / n - 320.0) < 320.0 * 1e-3


def test_rgb_round_trip():
    L = _lib()
    nm, d, rgb = np.zeros(4), np.zeros(4), np.zeros(3)
    rng = np.random.default_rng(2)
    for _ in range(10):
        c = rng.uniform(0.05, 0.95, 3)
        acc = np.zeros(3)
        n = 2048
        for i in range(n):
            L.or_sampled_wavelengths((i + 0.5) / n, O.ptr(nm), O.ptr(d))
            s = np.zeros(4)
            L.or_spectrum_of(O.ptr(c), O.ptr(nm), O.ptr(s))
            L.or_colour_of(O.ptr(nm), O.ptr(d), O.ptr(s), O.ptr(rgb))
            acc += rgb
        assert np.allclose(acc / n, c, atol=0.05)


def test_mixed_bits_is_splitmix64():
    from kodanuki_amd.scenes import mixed_bits
    for v in [0, 1, 2**63, 123456789]:
        assert O.lib().or_mixed_bits(v) == mixed_bits(v)
    # the first SplitMix64 outputs of seed 0 (published generator)
    assert O.lib().or_mixed_bits(0) == 0xE220A8397B1DCDAF


# ---- known answers of the reference's transport tests ----------------------
# (target/tests/transport.cpp:139-157 lit plane, :260-281 glass slab)

def lit_plane_text(light):
    return f"""import "standard.sdf";
render {{ width = 64; height = 40; seed = 1; samples = 256; surfaces = 2; far = 3; }}
camera {{ eye = [0, -3, 2]; target = [0, 0, 0]; up = [0, 0, 1]; fov = 0.3; }}
sky gradient {{ zenith = [0, 0, 0]; horizon = [0, 0, 0]; ground = [0, 0, 0]; }}
{light}
paint(matte([0.5, 0.5, 0.5])) at([0, 0, -0.1]) box([4, 4, 0.2]);
"""


QUAD = "light quad { at = [-0.1, -0.1, 1]; along = [0, 0.2, 0]; across = [0.2, 0, 0]; power = [1, 1, 1]; }"


def quad_irradiance(half, height, steps=400):
    """Irradiance under the centre of a square lamp of unit power (radiance
    1/area), by the midpoint rule: (1/area) * integral of h^2 / r^4."""
    d = 2 * half / steps
    c = -half + (np.arange(steps) + 0.5) * d
    x, y = np.meshgrid(c, c)
    r2 = x * x + y * y + height * height
    return float(np.sum(height * height / (r2 * r2)) * d * d / (4 * half * half))


def _window_mean(text, window, spp, depth=5):
    import kodanuki_amd as k
    s = k.Scene(text)
    cam, rs = s.settings()
    A = O.SceneArrays(s)
    lower, size = s.region(depth)
    x0, y0, x1, y1 = window
    img = O.render64(s, A, lower, size, rs.width, rs.height, spp, seed=rs.seed, surfaces=rs.surfaces, window=window, threads=8)
    assert img.shape == (y1 - y0, x1 - x0, 3)
    return img.mean(axis=(0, 1)) / spp


def test_lit_matte_plane_known_answer():
    """A grey matte plane under a quad lamp reads albedo/pi * irradiance
    (NEE, the lamp met along the bounce with MIS, radiance = power/area)."""
    got = _window_mean(lit_plane_text(QUAD), (31, 19, 33, 21), 1024)
    want = 0.5 / np.pi * quad_irradiance(0.1, 1.0)
    assert np.all(np.abs(got - want) < 0.02 * want), (got, want)


def slab_text(slab):
    return f"""import "standard.sdf";
render {{ width = 64; height = 40; seed = 1; samples = 256; surfaces = 7; far = 3; }}
camera {{ eye = [0, -0.3, 3]; target = [0, 0, 0]; up = [0, 0, 1]; fov = 0.3; }}
sky gradient {{ zenith = [0, 0, 0]; horizon = [0, 0, 0]; ground = [0, 0, 0]; }}
paint(matte([0, 0, 0], glow = [1, 1, 1], strength = 1)) at([0, 0, -0.1]) box([4, 4, 0.2]);
{slab}
"""


def test_glass_slab_known_answer():
    """A glowing plane seen through an exact glass slab keeps (1-R)/(1+R) of
    its radiance (the slab's transmission with every inner bounce)."""
    window = (22, 14, 42, 26)
    bare = _window_mean(slab_text(""), window, 64)
    through = _window_mean(slab_text("paint(glass(refraction = 1.5)) at([0, 0, 0.5]) box([3, 3, 0.4]);"), window, 64)
    cosine, ratio = 3 / np.sqrt(9.09), 1 / 1.5
    onward = np.sqrt(1 - ratio * ratio * (1 - cosine * cosine))
    perp = (ratio * cosine - onward) / (ratio * cosine + onward)
    para = (ratio * onward - cosine) / (ratio * onward + cosine)
    R = 0.5 * (perp * perp + para * para)
    want = (1 - R) / (1 + R)
    assert np.all(np.abs(through / bare - want) < 0.01 * want), (through / bare, want)
