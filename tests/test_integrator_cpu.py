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
