"""Adaptive stopping on the device (stopping/confidence_interval), through
the C ABI:
- the device's judge of every pixel's trial and its allotment from the
  neighbours' trials equal the oracle's (oracle/stopping_ref.py, restating
  confidence_interval.cpp): the splits and wanted counts to float rounding,
  the allotted counts exactly;
- the count pass draws exactly each pixel's allotment (its trial and the
  rest), in rounds small enough to split pixels between them too;
- a frame stopped by the confidence interval keeps the mean of a fixed
  count and draws fewer samples where pixels are quiet (the reference's
  checkStoppedMeans, target/tests/stopping.cpp:159-194: within 4 standard
  errors over 8 seeds, fewer than the ceiling of 64 and more than 16)."""
import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes
from oracle import stopping_ref as T
from tests.conftest import EMULATION

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)


def _setup(device, **kw):
    s = k.Scene(scenes.stopping_table(**kw))
    cam, rs = s.settings()
    lower, size = s.region(6)
    device.upload(s)
    device.build_octree(lower, size, 6)
    return s, rs, k.stopping_of(rs)


@pytest.mark.parametrize("flags", [0, k.PASS_SOBOL])
def test_allotment_equals_the_oracle(device, flags):
    s, rs, st = _setup(device)
    W, H = rs.width, rs.height
    trial = k.stopping_trial(st)
    assert trial == 16
    device.render_pass(W, H, trial, seed=rs.seed, surfaces=rs.surfaces, flags=k.PASS_ESTIMATES | flags, batch=st.batch)
    est = device.estimates(W, H)
    total = device.stopping_allot(st, trial)
    wanted, spreads, allotted = device.stopping_state(W, H)
    ow, osp, oa = T.allot_frame(est, st.samples, st.batch, st.tolerance, st.exposure, st.curve, trial)
    finite = np.isfinite(ow)
    assert np.array_equal(np.isfinite(wanted), finite)
    rel = np.abs(wanted[finite] - ow[finite]) / np.maximum(ow[finite], 1e-30)
    note(f"[stopping] wanted: {np.mean(wanted == ow):.4f} equal, worst relative gap {rel.max() if rel.size else 0:.2e}; "
         f"allotted mean {allotted.mean():.1f}, paths {total}")
    assert rel.size == 0 or rel.max() < 1e-6
    assert np.array_equal(spreads, osp)
    assert np.array_equal(allotted, oa)
    assert total == int((allotted - trial).sum())
    assert np.all(allotted % st.batch == 0) and allotted.min() >= trial and allotted.max() <= st.samples


@pytest.mark.parametrize("round_paths", [0, 4096 + 61])
def test_count_pass_draws_each_allotment(device, round_paths):
    s, rs, st = _setup(device)
    W, H = rs.width, rs.height
    trial, total, stats = device.render_stopped(W, H, st, seed=rs.seed, surfaces=rs.surfaces, round_paths=round_paths)
    _, _, allotted = device.stopping_state(W, H)
    rgb, sq, n = device.accumulation(W, H)
    est = device.estimates(W, H)
    note(f"[stopping] rounds of {round_paths or 'all'} paths: {total} paths after the trial, mean count {n.mean():.2f}")
    assert np.array_equal(n, allotted.astype(np.uint32))
    assert np.array_equal(est["taken"], allotted)
    assert np.allclose(est["total"], rgb, rtol=1e-12, atol=1e-12)
    # the count pass's estimates restart from the trial's totals (sumStart)
    assert np.all(est["batches"] == (allotted - trial) // st.batch)


def test_split_rounds_change_nothing(device):
    """Rounds that split pixels draw the same samples: the estimates, folded
    in sample order, equal to the bit; the accumulators, summed a round at a
    time, within rounding."""
    s, rs, st = _setup(device)
    W, H = rs.width, rs.height
    device.render_stopped(W, H, st, seed=rs.seed, surfaces=rs.surfaces)
    a, ea = device.accumulation(W, H), device.estimates(W, H)
    device.render_stopped(W, H, st, seed=rs.seed, surfaces=rs.surfaces, round_paths=5000)
    b, eb = device.accumulation(W, H), device.estimates(W, H)
    assert ea.tobytes() == eb.tobytes()
    assert np.array_equal(a[2], b[2])
    assert np.allclose(a[0], b[0], rtol=1e-12, atol=0) and np.allclose(a[1], b[1], rtol=1e-12, atol=0)


@pytest.mark.skipif(bool(EMULATION), reason="24 frames of 64x40: device only")
def test_confidence_interval_keeps_the_mean(device):
    s, rs, st = _setup(device)
    W, H = rs.width, rs.height
    fixed, stopped, counts = [], [], []
    for seed in range(1, 9):
        device.render_pass(W, H, st.samples, seed=seed, surfaces=rs.surfaces)
        rgb, _, n = device.accumulation(W, H)
        fixed.append((rgb / n[..., None]).mean(axis=(0, 1)))
        device.render_stopped(W, H, st, seed=seed, surfaces=rs.surfaces)
        rgb, _, n = device.accumulation(W, H)
        stopped.append((rgb / n[..., None]).mean(axis=(0, 1)))
        counts.append(n.mean())
    a, b = np.array(fixed), np.array(stopped)
    se = np.sqrt(a.var(axis=0, ddof=1) / len(a) + b.var(axis=0, ddof=1) / len(b))
    gap = np.max(np.abs(a.mean(axis=0) - b.mean(axis=0)) / np.maximum(se, 1e-12))
    note(f"[stopping] fixed {a.mean(axis=0)} at {st.samples}, confidence {b.mean(axis=0)} at {np.mean(counts):.1f} samples, "
         f"worst gap {gap:.2f} standard errors")
    assert gap < 4
    assert 16 < np.mean(counts) < 64
