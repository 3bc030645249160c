"""Adaptive stopping in plain Python: TEST INFRASTRUCTURE ONLY.

Only tests/ imports this module, to check the device's judge and allotment;
the product never does.

Restates the reference's stopping core and its confidence-interval member:
effortFor (core/stopping.cpp:18-34), splitSpread and splitCeiling
(core/stopping.cpp:85-108), displayed and stepsBetween
(shared/image.cpp:25-50), and confidenceTrial, confidenceWanted,
confidenceAllotted and confidenceSplit
(plugins/stopping/confidence_interval/confidence_interval.cpp:26-225), with
no prepared light (its spread zero). Pixel loops in Python: small images.

Pins (tests/test_stopping.py): effortFor, splitCeiling, splitSpread,
stepsBetween and the estimate fold against the reference's own functions
(oracle/_ref, built from its core/stopping.cpp and shared/image.cpp) and
fixtures written from them; the room's bisection against a root of the
graded span found independently; the allotment against the reference
test's properties (target/tests/stopping.cpp:196-254: a count is whole
batches set by the neighbours' trials alone) and closed forms.
"""
import math

import numpy as np

CONFIDENCE = 1.96
SMALLEST = 1e-30
LARGEST = 1e30
REFINEMENTS = 16
SHARED_COST = 1.5
TRIAL_BATCHES = 2
MOST_SPLIT = 8
FILMIC, CLAMPED = 0, 1


def effort_for(tolerance):
    """effortFor: (samples, batch)."""
    loosest, tightest, least, most, deepest, room = 62.5, 0.5, 8.0, 4096.0, 8.0, 4.0
    most_effort = math.pow(2.0, 0.66)
    if tolerance > 0:
        effort = math.log(loosest / tolerance) / math.log(loosest / tightest)
        effort = min(max(effort, 0.0), most_effort)
    else:
        effort = 1.0
    stride = int(_lround(1 + (deepest - 1) * effort))
    batch = (stride * stride + MOST_SPLIT - 1) // MOST_SPLIT * MOST_SPLIT
    ceiling = int(_lround((room if tolerance > 0 else 1.0) * least * math.pow(most / least, effort)))
    return max(batch, ceiling), batch


def _lround(v):
    """C's lround: halves away from zero."""
    return math.floor(v + 0.5) if v >= 0 else -math.floor(-v + 0.5)


def split_ceiling(batch):
    split = 1
    while split < MOST_SPLIT and batch > 0 and batch % (split * 2) == 0:
        split *= 2
    return split


def trial_samples(samples, batch, tolerance):
    """confidenceTrial."""
    n = TRIAL_BATCHES * batch
    return n if tolerance > 0 and n < samples else 0


def displayed(value, exposure, curve):
    if not value > 0:
        return 0.0
    exposed = value * float(np.float32(exposure))
    if curve == FILMIC:
        mapped = 2.51 / 2.43 if exposed >= 16.0 else exposed * (2.51 * exposed + 0.03) / (exposed * (2.43 * exposed + 0.59) + 0.14)
    else:
        mapped = exposed
    bounded = 0.0 if not mapped > 0 else (mapped if mapped < 1 else 1.0)
    return math.pow(bounded, 1 / 2.2)


def steps_between(low, high, exposure, curve):
    return abs(displayed(high, exposure, curve) - displayed(low, exposure, curve)) * 255.0


def _fits(mean, bar, exposure, curve, width):
    return steps_between(mean - width, mean + width, exposure, curve) <= bar


def _bisected(mean, bar, exposure, curve, low, high):
    for _ in range(REFINEMENTS):
        middle = math.sqrt(low * high)
        if _fits(mean, bar, exposure, curve, middle):
            low = middle
        else:
            high = middle
    return low


def room_around(mean, bar, exposure, curve):
    """The widest half width around mean whose graded span stays within bar."""
    start = max(abs(mean), SMALLEST)
    low = high = start
    if _fits(mean, bar, exposure, curve, start):
        while _fits(mean, bar, exposure, curve, high):
            if high > LARGEST:
                return math.inf
            low, high = high, high * 2
    else:
        while not _fits(mean, bar, exposure, curve, low):
            if low < SMALLEST:
                return 0.0
            high, low = low, low * 0.5
    return _bisected(mean, bar, exposure, curve, low, high)


def means_spread(means, squares, batches):
    count = float(batches)
    mean = means / count
    return max(squares / count - mean * mean, 0.0) * (count / (count - 1))


def batches_needed(mean, spread, tolerance, exposure, curve, unsure=0.0):
    if not spread > 0:
        return 0.0
    bar = 2 * tolerance
    whole = room_around(mean, bar, exposure, curve)
    theirs = CONFIDENCE * math.sqrt(unsure)
    own = whole
    if theirs > 0:
        alone = steps_between(mean - theirs, mean + theirs, exposure, curve) > bar
        own = whole if alone else math.sqrt(max(whole * whole - theirs * theirs, 0.0))
    if not own > 0:
        return math.inf
    return spread * (CONFIDENCE / own) * (CONFIDENCE / own)


def wanted(estimate, batch, tolerance, exposure, curve):
    """confidenceWanted of one estimate (a row of ESTIMATE_DTYPE), as float32."""
    batches = 0.0
    for c in range(3):
        mean = float(estimate["total"][c]) * (1.0 / int(estimate["taken"]))
        spread = means_spread(float(estimate["means"][c]), float(estimate["squares"][c]), int(estimate["batches"]))
        needed = batches_needed(mean, spread, tolerance, exposure, curve)
        batches = needed if batches < needed else batches
    return np.float32(batches * batch)


def split_spread(estimate, batch, split):
    """splitSpread: (shared, own) as float32."""
    if int(estimate["batches"]) < 2:
        return np.float32(0), np.float32(0)
    m = estimate["means"]
    own = (float(m[0]) + float(m[1]) + float(m[2])) / 3 - float(estimate["shared_means"])
    groups = float(batch) / float(split)
    b = int(estimate["batches"])
    return (np.float32(means_spread(float(estimate["shared_means"]), float(estimate["shared_squares"]), b) * groups),
            np.float32(means_spread(own, float(estimate["own_squares"]), b) * float(batch)))


def _split_near(ideal, ceiling):
    split = 1
    while split < ceiling and split * math.sqrt(2.0) <= ideal:
        split *= 2
    return split


def allotted(wanted_grid, spreads_grid, x, y, samples, batch, trial, trial_split):
    """confidenceAllotted and confidenceSplit of pixel (x, y) from its
    neighbours' trials: (allotted samples, the member's split)."""
    h, w = wanted_grid.shape
    total = shared = own = 0.0
    count = 0
    for row in range(max(y - 1, 0), min(y + 1, h - 1) + 1):
        for column in range(max(x - 1, 0), min(x + 1, w - 1) + 1):
            if row == y and column == x:
                continue
            total += float(wanted_grid[row, column])
            shared += float(spreads_grid[row, column, 0])
            own += float(spreads_grid[row, column, 1])
            count += 1
    s_shared = float(np.float32(shared / max(count, 1)))
    s_own = float(np.float32(own / max(count, 1)))
    ceiling = split_ceiling(batch)
    chosen = ceiling if not s_shared > 0 else _split_near(math.sqrt(s_own * SHARED_COST / s_shared), ceiling)
    if count == 0:
        return samples, chosen
    tried = s_shared * trial_split + s_own
    noise = 1.0 if not tried > 0 else (s_shared * chosen + s_own) / tried
    want = total / count * noise
    batches = math.ceil(math.sqrt(want * samples) / batch) if math.isfinite(want) else math.inf
    if not batches * batch < samples:
        return samples, chosen
    return max(int(batches) * batch, trial), chosen


def allot_frame(estimates, samples, batch, tolerance, exposure, curve, trial, split=1):
    """Every pixel's judged trial and allotment: (wanted [h, w] f32,
    spreads [h, w, 2] f32, allotted [h, w] i32)."""
    h, w = estimates.shape
    want = np.zeros((h, w), np.float32)
    spreads = np.zeros((h, w, 2), np.float32)
    for y in range(h):
        for x in range(w):
            e = estimates[y, x]
            want[y, x] = wanted(e, batch, tolerance, exposure, curve)
            spreads[y, x] = split_spread(e, batch, split)
    out = np.zeros((h, w), np.int32)
    for y in range(h):
        for x in range(w):
            out[y, x] = allotted(want, spreads, x, y, samples, batch, trial, split)[0]
    return want, spreads, out
