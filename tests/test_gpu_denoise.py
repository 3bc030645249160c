"""The denoiser on the device, through the C ABI: the same bits as the
oracle's restatement of the reference filter (oracle/denoise_ref.py) on the
reference's own test fields and on rendered histories, and the reference's
behavioural checks on the device's results."""
import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes
from oracle import denoise_ref as D
from tests import denoise_fields as F
from tests.conftest import EMULATION

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)


FIELDS = {
    "noisy_plane": lambda: (F.noisy_plane(), False),
    "edge_depth": lambda: (F.edge_field(False), False),
    "edge_tilted": lambda: (F.edge_field(True), False),
    "checker": lambda: (F.checker_kept_field(), False),
    "checker_demodulated": lambda: (F.checker_field(True), True),
    "converged": lambda: (F.converged_field(), False),
    "sky": lambda: (F.sky_field(), False),
    "ragged": lambda: (F.noisy_plane(37, 23), False),
}


@pytest.mark.parametrize("name", list(FIELDS))
def test_denoiser_matches_the_oracle_to_the_bit(device, name):
    px, demod = FIELDS[name]()
    h, w = px.shape
    got = device.denoise(px, F.FORWARD, F.RIGHT, F.UP, F.REACH, demodulated=demod)
    want = D.denoise(px, D.lens(F.FORWARD, F.RIGHT, F.UP, F.REACH, w, h), demod)
    diff = np.abs(got.astype(np.float64) - want)
    note(f"[denoise] {name} {w}x{h}: max |device - oracle| {diff.max():.3g}, equal bits {np.mean(got.view(np.uint32) == want.view(np.uint32)):.6f}")
    assert got.tobytes() == want.tobytes()


def test_device_passes_the_reference_checks(device):
    px = F.noisy_plane()
    img = device.denoise(px, F.FORWARD, F.RIGHT, F.UP, F.REACH)
    assert F.red_error(img, 0.5) < F.red_error(F.colour_image(px), 0.5) / 5
    for tilted in (False, True):
        img = device.denoise(F.edge_field(tilted), F.FORWARD, F.RIGHT, F.UP, F.REACH)
        assert abs(img[:, F.EDGE - 1, 0].mean() - 0.4) < 0.03 and abs(img[:, F.EDGE, 0].mean() - 0.6) < 0.03
    light = device.denoise(F.checker_field(False), F.FORWARD, F.RIGHT, F.UP, F.REACH)
    divided = device.denoise(F.checker_field(True), F.FORWARD, F.RIGHT, F.UP, F.REACH, demodulated=True)
    assert np.max(np.abs(divided.astype(np.float64) - light)) < 1e-4


def test_denoised_render_is_closer_to_the_converged_image(device):
    """A 4-sample render denoised from its own accumulators is nearer the
    converged image than the raw 4-sample mean (the filter is biased: it
    trades detail for less noise; at 4 samples a pixel shows 2/3 of the
    filtered light)."""
    text = scenes.plain(0, width=96, height=64)
    s = k.Scene(text)
    cam, rs = s.settings()
    lower, size = s.region(6)
    device.upload(s)
    device.build_octree(lower, size, 6)
    passes = 2 if EMULATION else 8
    for p in range(passes):
        device.render_pass(96, 64, 64, first_sample=(1 << 20) + p * 64, seed=rs.seed, surfaces=rs.surfaces, keep=p > 0)
    rgb, _, n = device.accumulation(96, 64)
    ref = rgb / n[..., None]
    device.render_pass(96, 64, 4, seed=rs.seed, surfaces=rs.surfaces)
    rgb, _, n = device.accumulation(96, 64)
    raw = rgb / n[..., None]
    den = device.denoise_accumulation(96, 64)
    assert np.all(np.isfinite(den))

    def rel(img):
        return float(np.mean((img - ref) ** 2 / (ref ** 2 + 1e-2)))

    note(f"[denoise] plain scene 96x64 at 4 spp: relMSE raw {rel(raw):.4e}, denoised {rel(den):.4e} "
         f"(reference {64 * passes} spp)")
    # a third of the raw error at least (the filtered share at 4 samples is 2/3)
    assert rel(den) < rel(raw) / 1.3
