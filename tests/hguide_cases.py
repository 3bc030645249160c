"""The plain guide's test lumps and sphere bins, as the reference's own
sampler tests lay them out (target/tests/sampler.cpp:438-453 binDirection
and binOf over 32 x 64 bins of equal area; :588-625 testLumps: a narrow
lump toward (0.3, 0.1, 0.9) of weight 2 and a wide one toward
(-0.5, 0.2, 0.7) of weight 1, means (0.03, -0.02), points two units along
their axes, seen from (0.1, -0.2, 0); lumpOutgoing (0.2, -0.3, 0.9)
normalised; normal +z)."""
import numpy as np

from oracle import hguide_ref as H

ROWS, COLUMNS = 32, 64
NORMAL = np.array([0.0, 0.0, 1.0])
OUTGOING = np.array([0.2, -0.3, 0.9]) / np.linalg.norm([0.2, -0.3, 0.9])
PLACE = np.array([0.1, -0.2, 0.0])


def _basis(axis):
    helper = np.array([1.0, 0, 0]) if abs(axis[0]) < 0.9 else np.array([0, 1.0, 0])
    t = np.cross(helper, axis)
    return t / np.linalg.norm(t)


def _lump(toward, factor, weight):
    axis = np.asarray(toward, float) / np.linalg.norm(toward)
    return dict(axis=axis, across=_basis(axis), point=axis * 2, mean=(0.03, -0.02), factor=factor, weight=weight)


def test_guide():
    return H.pack([_lump((0.3, 0.1, 0.9), (0.08, 0.02, 0.06), 2), _lump((-0.5, 0.2, 0.7), (0.15, 0, 0.1), 1)], share=0.5)


def bin_directions(steps):
    """[bins, steps^2, 3]: the midpoints of a steps x steps grid in each bin."""
    row = np.arange(ROWS * COLUMNS) // COLUMNS
    col = np.arange(ROWS * COLUMNS) % COLUMNS
    k = np.arange(steps * steps)
    wu, wv = (k % steps + 0.5) / steps, (k // steps + 0.5) / steps
    height = -1 + 2 * (row[:, None] + wu[None, :]) / ROWS
    angle = 2 * np.pi * (col[:, None] + wv[None, :]) / COLUMNS
    ring = np.sqrt(np.maximum(1 - height * height, 0))
    return np.stack([ring * np.cos(angle), ring * np.sin(angle), height], axis=-1)


def bin_of(d):
    angle = np.arctan2(d[:, 1], d[:, 0])
    turn = np.where(angle < 0, angle + 2 * np.pi, angle)
    row = np.minimum(((d[:, 2] + 1) / 2 * ROWS).astype(int), ROWS - 1)
    col = np.minimum((turn / (2 * np.pi) * COLUMNS).astype(int), COLUMNS - 1)
    return row * COLUMNS + col


def bin_masses(density_of, steps=32):
    """Each bin's mass under a density of directions (midpoint rule)."""
    d = bin_directions(steps)
    dens = density_of(d.reshape(-1, 3)).reshape(ROWS * COLUMNS, steps * steps)
    return dens.mean(axis=1) * (4 * np.pi / (ROWS * COLUMNS))


def chi_squared(counts, expected):
    """The reference's statistic: bins expecting under 5 pooled; the bound
    degrees + 3.09 sqrt(2 degrees) + 10 (sampler.cpp:680-690)."""
    small = expected < 5
    chi = np.sum((counts[~small] - expected[~small]) ** 2 / expected[~small])
    pooled = (counts[small].sum(), expected[small].sum())
    chi += (pooled[0] - pooled[1]) ** 2 / pooled[1] if pooled[1] > 0 else pooled[0]
    degrees = int(np.sum(~small))
    return chi, degrees, degrees + 3.09 * np.sqrt(2.0 * degrees) + 10


def above_both(d):
    return (d @ NORMAL > 0) & (OUTGOING @ NORMAL > 0)
