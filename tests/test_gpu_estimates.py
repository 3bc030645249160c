"""Per-pixel estimates on the device (the reference's Estimate,
core/stopping.h:57-100), through the C ABI:
- the device's fold of given sample colours equals the oracle's
  (oracle/estimate_ref.py) to the bit, ragged batches and a second call
  continuing the first included;
- a render pass that keeps estimates (flags bit 6) counts every sample,
  closes a batch every `batch` samples across progressive passes, and its
  sample-order totals agree with the accumulators' sums."""
import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes
from oracle import estimate_ref as E

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("batch", [0, 1, 3, 8])
def test_device_fold_equals_the_oracle(device, batch):
    rng = np.random.default_rng(10 + batch)
    pixels = 37
    first = rng.lognormal(-1, 1.5, (pixels, 10, 3))
    second = rng.lognormal(-1, 1.5, (pixels, 7, 3))
    got = device.fold_samples(second, batch, device.fold_samples(first, batch))
    for p in range(pixels):
        want = E.fold_pixel(E.fold_pixel(E.empty(), first[p], batch), second[p], batch)
        assert got[p].tobytes() == E.as_record(want, k.ESTIMATE_DTYPE).tobytes(), (batch, p, got[p], want)


def test_render_pass_keeps_estimates(device):
    s = k.Scene(scenes.tiny(glass=False))
    cam, rs = s.settings()
    lower, size = s.region(5)
    device.upload(s)
    device.build_octree(lower, size, 5)
    W, H = 40, 24
    flags = k.PASS_ESTIMATES
    for batch in (1, 4):
        device.render_pass(W, H, 6, seed=rs.seed, surfaces=4, flags=flags, batch=batch)
        device.render_pass(W, H, 10, first_sample=6, seed=rs.seed, surfaces=4, keep=True, flags=flags, batch=batch)
        est = device.estimates(W, H)
        rgb, sq, n = device.accumulation(W, H)
        assert np.all(est["taken"] == 16) and np.all(n == 16)
        assert np.all(est["batches"] == 16 // batch)
        # the same samples summed in sample order (the estimate) and by
        # slices (the accumulators)
        assert np.allclose(est["total"], rgb, rtol=1e-12, atol=1e-12)
        if batch == 1:
            assert est["means"].tobytes() == est["total"].tobytes()
        else:
            assert np.allclose(est["means"] / est["batches"][..., None], est["total"] / 16, rtol=1e-12, atol=1e-14)
        note(f"[estimates] batch {batch}: mean of the batch means {est['means'].mean() / (16 // batch):.6f}")
