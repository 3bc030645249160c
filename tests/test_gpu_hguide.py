"""The plain (highlight) guide on the device, through the C ABI: its lump
density and draws against the oracle's restatement of the reference's
(oracle/hguide_ref.py) and the reference's chi-squared check
(target/tests/sampler.cpp:646-690); the survey's guides of a room whose
far wall is lit off a steel plate; a guided render's mean unchanged
within 4 standard errors of the plain render's (the reference's pairing
check, sampler.cpp:129-171), and its error at equal samples lower."""
import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes
from oracle import hguide_ref as H
from tests import hguide_cases as T
from tests.conftest import EMULATION

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)


def test_lump_density_matches_the_oracle(device):
    g = T.test_guide()
    rng = np.random.default_rng(5)
    d = rng.normal(size=(4096, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    # and directions near the lumps, where the density is large
    for L in H.unpack(g):
        near = L["axis"][None, :] + rng.normal(scale=0.1, size=(1024, 3))
        d = np.concatenate([d, near / np.linalg.norm(near, axis=1)[:, None]])
    d = d.astype(np.float32)
    got = device.highlight_density(g, T.PLACE, T.NORMAL, T.OUTGOING, d).astype(np.float64)
    want = H.density(g, T.PLACE, T.NORMAL, T.OUTGOING, d.astype(np.float64))
    err = np.abs(got - want)
    note(f"[hguide] density: max {want.max():.4g}, largest |device - oracle| {err.max():.3g}")
    assert np.all(err <= 1e-4 * want + 1e-6 * want.max())


def test_lump_draws_match_the_oracle(device):
    g = T.test_guide()
    q = np.random.default_rng(6).random((8192, 3)).astype(np.float32)
    got = device.highlight_draw(g, T.PLACE, q).astype(np.float64)
    want = H.draw(g, T.PLACE, q.astype(np.float64))
    err = np.abs(got - want).max()
    note(f"[hguide] draws: largest |device - oracle| {err:.3g}")
    assert err < 2e-5


def test_device_lumps_draw_by_their_density(device):
    g = T.test_guide()
    draws = 100_000 if EMULATION else 1_000_000
    d = device.highlight_draw(g, T.PLACE, np.random.default_rng(11).random((draws, 3)).astype(np.float32)).astype(np.float64)
    found = T.above_both(d)
    counts = np.bincount(T.bin_of(d[found]), minlength=T.ROWS * T.COLUMNS).astype(float)
    masses = T.bin_masses(lambda x: device.highlight_density(g, T.PLACE, T.NORMAL, T.OUTGOING, x.astype(np.float32)).astype(np.float64),
                          steps=16)
    chi, degrees, bound = T.chi_squared(counts, masses * draws)
    note(f"[hguide] device chi squared {chi:.1f} over {degrees} bins (bound {bound:.1f}), integral {masses.sum():.4f}, "
         f"found {found.mean():.4f}")
    assert degrees > 20 and chi < bound
    assert abs(masses.sum() - found.mean()) < 0.01


def test_bad_guides_are_refused(device):
    g = T.test_guide()
    for bad in (0, 13):
        h = g.copy()
        h[0] = bad
        with pytest.raises(k.KodanukiError):
            device.highlight_density(h, T.PLACE, T.NORMAL, T.OUTGOING, np.zeros((1, 3), np.float32))
    h = g.copy()
    h[H.HEAD + 14] = -1
    with pytest.raises(k.KodanukiError):
        device.highlight_draw(h, T.PLACE, np.zeros((1, 3), np.float32))


W, H_ = (40, 30) if EMULATION else (96, 72)


@pytest.fixture(scope="module")
def room(device):
    s = k.Scene(scenes.highlight_room(W, H_))
    lower, size = s.region(6)
    device.upload(s)
    device.build_octree(lower, size, 6)
    st = device.highlight_guides(W, H_)
    note(f"[hguide] survey of {W}x{H_}: {st.survey_points} points, {st.highlights} highlights, {st.guided_pixels} guided pixels, "
         f"{st.mean_lumps:.2f} lumps, share {st.mean_share:.3f}, {st.ms:.2f} ms")
    return s, st


def test_survey_guides_the_wall_toward_the_plate(device, room):
    s, st = room
    assert st.highlights > 0 and st.guided_pixels > W * H_ // 10
    guided = 0
    for y in range(0, H_, 3):
        for x in range(0, W, 3):
            g = device.highlight_guide_at(x, y)
            n = int(g[0])
            if n == 0:
                continue
            guided += 1
            assert 1 <= n <= 12 and 0.02 <= g[2] <= 0.5
            lumps = g[H.HEAD:H.HEAD + n * H.LUMP_FLOATS].reshape(n, H.LUMP_FLOATS)
            assert np.all(lumps[:, 14] > 0)
            assert np.allclose(np.linalg.norm(lumps[:, 0:3], axis=1), 1, atol=1e-5)
            # every lump's point lies on the plate's plane, within 0.1 of the
            # plate (a highlight walked again from the pixel moves at most 0.1)
            assert np.all(np.abs(lumps[:, 6:8]) <= 0.71) and np.all(np.abs(lumps[:, 8] - 0.02) < 0.01)
    assert guided > 0


def _frame(device, spp, seed, flags):
    device.render_pass(W, H_, spp, seed=seed, surfaces=6, flags=flags)
    rgb, _, n = device.accumulation(W, H_)
    return rgb / n[..., None]


def test_guided_mean_is_unchanged(device, room):
    seeds, spp = (3, 4) if EMULATION else (12, 64)
    arms = {}
    for name, flags in (("plain", 0), ("guided", k.PASS_HIGHLIGHT)):
        means = np.array([_frame(device, spp, 100 + i, flags).mean(axis=(0, 1)) for i in range(seeds)])
        arms[name] = (means.mean(axis=0), means.std(axis=0, ddof=1) / np.sqrt(seeds))
    (a, ea), (b, eb) = arms["plain"], arms["guided"]
    gap = np.abs(a - b) / np.maximum(np.sqrt(ea * ea + eb * eb), 1e-12)
    note(f"[hguide] plain {a} (error {ea[1]:.2e}), guided {b} (error {eb[1]:.2e}), worst gap {gap.max():.2f} standard errors")
    assert gap.max() < 4


@pytest.mark.skipif(bool(EMULATION), reason="a converged reference takes minutes on the CPU")
def test_guided_error_is_lower_at_equal_samples(device, room):
    reference = np.mean([_frame(device, 1024, 900 + i, 0) for i in range(4)], axis=0)
    lum = lambda im: im @ np.array([0.2126, 0.7152, 0.0722])
    mask = lum(reference) > 0
    errs = {}
    for name, flags in (("plain", 0), ("guided", k.PASS_HIGHLIGHT)):
        e = [np.mean(((lum(_frame(device, 16, 10 + i, flags)) - lum(reference))[mask] / np.maximum(lum(reference)[mask], 1e-3)) ** 2)
             for i in range(4)]
        errs[name] = float(np.mean(e))
    gain = errs["plain"] / errs["guided"]
    note(f"[hguide] relative MSE at 16 spp: plain {errs['plain']:.4g}, guided {errs['guided']:.4g}, gain {gain:.2f}x")
    assert gain > 1.2


def test_sphere_lamp_is_surveyed_and_the_mean_kept(device):
    """The survey mirrors sphere lamps as it mirrors quads; a guided render
    of the room lit by a ball keeps its mean within 4 standard errors."""
    s = k.Scene(scenes.highlight_room(W, H_, lamp="sphere"))
    lower, size = s.region(6)
    device.upload(s)
    device.build_octree(lower, size, 6)
    st = device.highlight_guides(W, H_)
    note(f"[hguide] sphere lamp: {st.highlights} highlights, {st.guided_pixels} guided pixels")
    assert st.highlights > 0 and st.guided_pixels > 0
    seeds, spp = (3, 4) if EMULATION else (12, 64)
    arms = {}
    for name, flags in (("plain", 0), ("guided", k.PASS_HIGHLIGHT)):
        means = np.array([_frame(device, spp, 200 + i, flags).mean(axis=(0, 1)) for i in range(seeds)])
        arms[name] = (means.mean(axis=0), means.std(axis=0, ddof=1) / np.sqrt(seeds))
    (a, ea), (b, eb) = arms["plain"], arms["guided"]
    gap = np.abs(a - b) / np.maximum(np.sqrt(ea * ea + eb * eb), 1e-12)
    note(f"[hguide] sphere lamp: plain {a}, guided {b}, worst gap {gap.max():.2f} standard errors")
    assert gap.max() < 4
