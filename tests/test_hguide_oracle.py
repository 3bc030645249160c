"""The plain guide's CPU oracle (oracle/hguide_ref.py) against the
reference's own lump checks (target/tests/sampler.cpp:646-690
checkLumpDensity: draws follow the density by a chi-squared test over
32 x 64 bins, and the density's integral is the share of draws found) and
against closed forms: a lone round lump's peak is 1 / (2 pi s^2) over the
kept mass, its density integrates to one, and seen from anywhere on the
line through its point it peaks toward the point."""
import numpy as np

from oracle import hguide_ref as H
from tests import hguide_cases as T


def test_lumps_draw_by_the_density_they_report():
    g = T.test_guide()
    rng = np.random.default_rng(11)
    draws = 1_000_000
    d = H.draw(g, T.PLACE, rng.random((draws, 3)))
    found = T.above_both(d)
    counts = np.bincount(T.bin_of(d[found]), minlength=T.ROWS * T.COLUMNS).astype(float)
    masses = T.bin_masses(lambda x: H.density(g, T.PLACE, T.NORMAL, T.OUTGOING, x))
    chi, degrees, bound = T.chi_squared(counts, masses * draws)
    print(f"chi squared {chi:.1f} over {degrees} bins (bound {bound:.1f}), integral {masses.sum():.4f}, found {found.mean():.4f}")
    assert degrees > 20 and chi < bound
    assert abs(masses.sum() - found.mean()) < 0.01


def test_lumps_with_a_mistake_fail_the_chi_squared_test():
    """The test has teeth: a density that forgets the 1 / along^3 of the
    map from the plane to the sphere fails it."""
    g = T.test_guide()
    rng = np.random.default_rng(11)
    draws = 1_000_000
    d = H.draw(g, T.PLACE, rng.random((draws, 3)))
    found = T.above_both(d)
    counts = np.bincount(T.bin_of(d[found]), minlength=T.ROWS * T.COLUMNS).astype(float)

    def flat(x):
        lumps = H.unpack(g)
        total = np.zeros(len(x))
        for L in lumps:
            along = np.maximum(x @ L["axis"], 1e-9)
            total += L["weight"] * H.lump_at(L, H.centre(L, T.PLACE), x) * along ** 3
        return np.where(T.above_both(x), total, 0)
    chi, degrees, bound = T.chi_squared(counts, T.bin_masses(flat) * draws)
    assert chi > bound


def _round_lump(sigma, axis=(0, 0, 1.0), point=(0, 0, 2.0)):
    return H.pack([dict(axis=np.array(axis), across=np.array([1.0, 0, 0]), point=np.array(point), mean=(0, 0),
                        factor=(sigma, 0, sigma), weight=1.0)])


def test_round_lump_peak_and_integral():
    sigma = 0.05
    g = _round_lump(sigma)
    peak = H.density(g, [0, 0, 0], [0, 0, 1], [0, 0, 1], [[0, 0, 1]])[0]
    assert abs(peak / (1 / (2 * np.pi * sigma * sigma) / H.KEPT) - 1) < 1e-6
    # over the hemisphere by a fine grid in the lump's plane: exactly the
    # Gaussian over the plane
    u = np.linspace(-8 * sigma, 8 * sigma, 801)
    U, V = np.meshgrid(u, u)
    d = np.stack([U.ravel(), V.ravel(), np.ones(U.size)], axis=1)
    d /= np.linalg.norm(d, axis=1)[:, None]
    jac = 1 / d[:, 2] ** 3  # solid angle per plane area at distance one
    mass = np.sum(H.density(g, [0, 0, 0], [0, 0, 1], [0, 0, 1], d) / jac) * (u[1] - u[0]) ** 2
    assert abs(mass - 1) < 1e-6


def test_shifted_centre_aims_at_the_point():
    g = _round_lump(0.02, axis=(0, 0, 1.0), point=(0.0, 0.0, 2.0))
    for place in ([0.3, -0.1, 0.0], [-0.5, 0.4, 0.5]):
        to = np.array([0, 0, 2.0]) - place
        to /= np.linalg.norm(to)
        L = H.unpack(g)[0]
        c = H.centre(L, place)
        # the lump's peak direction from place is the point's direction
        peak = L["axis"] + L["across"] * c[0] + L["up"] * c[1]
        peak /= np.linalg.norm(peak)
        assert np.allclose(peak, to, atol=1e-6)
        d = H.draw(g, place, np.random.default_rng(3).random((20000, 3)))
        mean = d.mean(axis=0)
        assert np.dot(mean / np.linalg.norm(mean), to) > 1 - 1e-3


def test_pick_follows_the_weights():
    g = T.test_guide()
    pick = np.linspace(0, 1, 10001, endpoint=False)
    q = np.stack([pick, np.full_like(pick, 0.3), np.full_like(pick, 0.6)], axis=1)
    d = H.draw(g, T.PLACE, q)
    lumps = H.unpack(g)
    nearest = np.argmax(np.stack([d @ L["axis"] for L in lumps], axis=1), axis=1)
    assert abs(np.mean(nearest == 0) - 2 / 3) < 1e-3
