"""The stopping oracle (oracle/stopping_ref.py) against what fixes it.

effortFor, splitCeiling, splitSpread and stepsBetween are pinned to the
reference's own functions (tests/test_ref_pins.py). Here: the
confidence-interval member's judge and allotment, which the reference only
builds with its generated plugin registry, against the reference test's
properties (target/tests/stopping.cpp:196-254) and closed forms.
"""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

import kodanuki_amd as k
from oracle import estimate_ref as E
from oracle import stopping_ref as T


def test_own_trial_unread_and_whole_batches():
    """A pixel's count is whole batches set by its neighbours' trials alone,
    and its split divides a batch (checkOwnTrialUnread: 1024 samples,
    batches of 8, tolerance 2, trials of 16 drawn two to a first vertex)."""
    rng = np.random.default_rng(7)
    wanted = rng.uniform(0, 2000, (5, 7)).astype(np.float32)
    spreads = rng.uniform(0, 2000, (5, 7, 2)).astype(np.float32)
    for y in range(5):
        for x in range(7):
            own_w, own_s = wanted[y, x], spreads[y, x].copy()
            wanted[y, x], spreads[y, x] = 0, (0, 1e9)
            low, few = T.allotted(wanted, spreads, x, y, 1024, 8, 16, 2)
            wanted[y, x], spreads[y, x] = 1e9, (1e9, 0)
            high, many = T.allotted(wanted, spreads, x, y, 1024, 8, 16, 2)
            wanted[y, x], spreads[y, x] = own_w, own_s
            assert low == high and low % 8 == 0 and 16 <= low <= 1024
            assert few == many and 8 % few == 0


@pytest.mark.parametrize("curve", [T.FILMIC, T.CLAMPED])
@pytest.mark.parametrize("mean", [1e-4, 0.01, 0.18, 0.5, 0.9, 3.0])
def test_room_is_the_root_of_the_graded_span(mean, curve):
    """The room is the half width whose graded span reaches the bar, to the
    bisection's precision (16 geometric halvings of a factor of 2)."""
    bar = 2 * 2.0
    room = T.room_around(mean, bar, 1.0, curve)
    f = lambda w: T.steps_between(mean - w, mean + w, 1.0, curve) - bar  # noqa: E731
    if f(1e6) <= 0:
        assert room == math.inf
        return
    root = brentq(f, 1e-12, 1e6, xtol=1e-15, rtol=1e-14)
    assert f(room) <= 0
    assert root * (1 - 2e-5) <= room <= root * (1 + 1e-12)


def test_wanted_is_spread_over_room_squared():
    """confidenceWanted = batch x spread x (1.96 / room)^2 on the worst channel."""
    rng = np.random.default_rng(3)
    colours = np.exp(rng.normal(-1.5, 0.8, (16, 3)))
    e = E.as_record(E.fold_pixel(E.empty(), colours, 8), k.ESTIMATE_DTYPE)
    want = T.wanted(e, 8, 2.0, 1.0, T.FILMIC)
    worst = 0.0
    for c in range(3):
        mean = float(e["total"][c]) / int(e["taken"])
        bm = [colours[:8, c].sum() / 8, colours[8:, c].sum() / 8]
        spread = np.var(bm, ddof=1)
        room = T.room_around(mean, 4.0, 1.0, T.FILMIC)
        worst = max(worst, spread * (1.96 / room) ** 2)
    assert want == pytest.approx(worst * 8, rel=1e-6)


@pytest.mark.parametrize("w", [0.0, 3.0, 40.0, 500.0, 5000.0, np.inf])
def test_uniform_want_closed_form(w):
    """Where every neighbour wants w samples and the shared part carries all
    the spread: ceil(sqrt(w S) / B) B, at least the trial, the ceiling where
    that reaches it."""
    S, B, trial = 256, 8, 16
    wanted = np.full((4, 5), w, np.float32)
    spreads = np.zeros((4, 5, 2), np.float32)
    spreads[..., 0] = 1.0
    got = {T.allotted(wanted, spreads, x, y, S, B, trial, 1)[0] for y in range(4) for x in range(5)}
    if not np.isfinite(w):
        expect = S
    else:
        b = math.ceil(math.sqrt(w * S) / B)
        expect = S if not b * B < S else max(b * B, trial)
    assert got == {expect}


def test_a_lone_pixel_takes_the_ceiling():
    assert T.allotted(np.zeros((1, 1), np.float32), np.zeros((1, 1, 2), np.float32), 0, 0, 64, 8, 16, 1)[0] == 64


def test_trial_is_two_batches_below_the_ceiling():
    assert T.trial_samples(64, 8, 6.0) == 16
    assert T.trial_samples(16, 8, 6.0) == 0
    assert T.trial_samples(64, 8, 0.0) == 0
    s = k.Stopping(64, 8, 6.0, 1.0, k.CURVE_FILMIC, 1)
    assert k.stopping_trial(s) == 16
    with pytest.raises(k.KodanukiError):
        k.stopping_trial(k.Stopping(64, 8, 6.0, 1.0, k.CURVE_FILMIC, 2))
