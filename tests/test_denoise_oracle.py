"""The denoiser's CPU oracle against the reference's own behavioural checks
(target/tests/denoise.cpp:143-312), which hold no numeric golden: noise on
a plane flattened at least 5 times, the two sides of an edge (in depth or
in the normal) keep their light, a checkered albedo keeps its checks while
its light is filtered, a history of lighting shows what the same history
of light shows, a converged field comes out nearly unchanged; and the
quick exponential within 0.3% of exp."""
import numpy as np

from oracle import denoise_ref as D
from tests import denoise_fields as F


def _run(px, demodulated=False):
    h, w = px.shape
    return D.denoise(px, D.lens(F.FORWARD, F.RIGHT, F.UP, F.REACH, w, h), demodulated)


def test_quick_exp_is_close_to_exp():
    x = -np.linspace(0, 29.9, 20001).astype(np.float32)
    q = D.quick_exp(x).astype(np.float64)
    assert np.max(np.abs(q / np.exp(x.astype(np.float64)) - 1)) < 3e-3
    assert D.quick_exp(np.float32(-30.0)) == 0 and D.quick_exp(np.float32(-31)) == 0


def test_noise_on_a_plane_is_flattened():
    px = F.noisy_plane()
    before = F.red_error(F.colour_image(px), 0.5)
    after = F.red_error(_run(px), 0.5)
    assert after < before / 5


def test_edges_keep_their_sides():
    for tilted in (False, True):
        img = _run(F.edge_field(tilted))
        assert abs(img[:, F.EDGE - 1, 0].mean() - 0.4) < 0.03
        assert abs(img[:, F.EDGE, 0].mean() - 0.6) < 0.03


def test_checkered_albedo_keeps_its_checks():
    px = F.checker_kept_field()
    before = F.red_error(F.colour_image(px), F.checker_albedo)
    after = F.red_error(_run(px), F.checker_albedo)
    assert after < before / 4


def test_demodulated_history_shows_the_same():
    light = _run(F.checker_field(False))
    divided = _run(F.checker_field(True), demodulated=True)
    assert np.max(np.abs(divided.astype(np.float64) - light)) < 1e-4


def test_converged_field_is_nearly_unchanged():
    px = F.converged_field()
    assert np.max(np.abs(_run(px).astype(np.float64) - F.colour_image(px))) < 0.01


def test_sky_is_filtered_with_the_sky_alone():
    px = F.sky_field()
    img = _run(px)
    # the sky's light (about 1.0) does not leak into the plane (about 0.5)
    assert abs(img[F.HEIGHT // 2:, :, 0].mean() - 0.5) < 0.03
    assert abs(img[:F.HEIGHT // 2, :, 0].mean() - 1.0) < 0.05
