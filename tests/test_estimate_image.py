"""A pixel's estimate (the reference's Estimate, core/stopping.h:57-100) and
the image outputs (writePpm and writePfm, shared/image.cpp:55-96) on the
CPU: the oracle (oracle/estimate_ref.py) pinned by what the reference's
definitions fix in closed form, and the library's host image functions
against it through the C ABI (no GPU)."""
import math
import os

import numpy as np
import pytest

import kodanuki_amd as k
from oracle import estimate_ref as E


def test_fold_with_batches_of_one_sums_the_samples():
    """With a batch of one, every sample closes a batch: the means add the
    samples themselves in order (times the reciprocal 1.0), so they equal
    the total to the bit, and the squares the samples' squares."""
    rng = np.random.default_rng(1)
    cols = rng.lognormal(size=(50, 3))
    e = E.fold_pixel(E.empty(), cols, 1)
    assert e["taken"] == 50 and e["batches"] == 50 and e["batch"] == [0.0, 0.0, 0.0]
    assert e["means"] == e["total"]
    sq = [0.0, 0.0, 0.0]
    for c in cols:
        sq = [sq[k] + float(c[k]) * float(c[k]) for k in range(3)]
    assert e["squares"] == sq


def test_fold_of_exact_colours_closes_batches_by_brute_force():
    """Colours that are multiples of 1/64 below 4 add without rounding: the
    batch means and their squares are then what a brute-force reshape of
    the samples into batches gives, the last unclosed batch is held, and
    equal channels leave nothing of a batch's own part."""
    rng = np.random.default_rng(2)
    spp, batch = 14, 4
    grey = rng.integers(0, 256, spp) / 64.0
    cols = np.stack([grey, grey, grey], axis=1)
    e = E.fold_pixel(E.empty(), cols, batch)
    whole = spp // batch
    means = cols[:whole * batch].reshape(whole, batch, 3).sum(axis=1) / batch
    assert e["batches"] == whole and e["taken"] == spp
    assert e["means"] == means.sum(axis=0).tolist()
    assert e["squares"] == (means ** 2).sum(axis=0).tolist()
    assert e["batch"] == cols[whole * batch:].sum(axis=0).tolist()
    assert e["shared_means"] == float(means[:, 0].sum())
    assert e["own_squares"] == 0.0
    # a constant pixel has no spread between its batches (meansSpread,
    # core/stopping.cpp:86-90)
    c = E.fold_pixel(E.empty(), np.full((16, 3), 0.375), 4)
    n = c["batches"]
    spread = max(c["squares"][0] / n - (c["means"][0] / n) ** 2, 0.0) * (n / (n - 1))
    assert spread == 0.0


def test_fold_of_no_batch_only_sums():
    e = E.fold_pixel(E.empty(), np.ones((5, 3)), 0)
    assert e["batches"] == 0 and e["means"] == [0.0, 0.0, 0.0] and e["batch"] == [5.0, 5.0, 5.0]


def test_grade_closed_forms():
    # nothing, negative and NaN light show black; the filmic curve
    # saturates at 16 (its ceiling 2.51 / 2.43 clamps to full scale)
    assert E.displayed(0.0) == 0.0 and E.displayed(-1.0) == 0.0 and E.displayed(float("nan")) == 0.0
    assert E.displayed(16.0) == 1.0 and E.displayed(1e9) == 1.0
    # the filmic curve at 1: 2.54 / 3.16, to the power 1/2.2
    assert abs(E.displayed(1.0) - math.pow(2.54 / 3.16, 1 / 2.2)) < 1e-15
    assert E.encode(np.float32([1.0]))[0] == round(255 * math.pow(2.54 / 3.16, 1 / 2.2))
    # clamped: full scale at 1 and beyond; a quarter shows 0.25^(1/2.2)
    assert E.encode(np.float32([1.0, 7.0]), curve=1).tolist() == [255, 255]
    assert E.encode(np.float32([0.25]), curve=1)[0] == round(255 * 0.25 ** (1 / 2.2))
    assert E.encode(np.float32([0.125]), exposure=2.0, curve=1)[0] == E.encode(np.float32([0.25]), curve=1)[0]


@pytest.mark.parametrize("curve", [k.CURVE_FILMIC, k.CURVE_CLAMPED])
def test_library_grades_as_the_oracle(curve):
    rng = np.random.default_rng(3)
    img = np.concatenate([rng.lognormal(-1, 1.5, 3 * 40 * 30), [0, -1, 16, 20, 1e30, np.nan]]).astype(np.float32)[:3 * 40 * 30]
    img = img.reshape(30, 40, 3)
    got = k.encode_image(img, exposure=1.5, curve=curve)
    assert got.tobytes() == E.encode(img, 1.5, curve).tobytes()


def test_ppm_and_pfm_files(tmp_path):
    rng = np.random.default_rng(4)
    img = rng.lognormal(-1, 1, (7, 5, 3)).astype(np.float32)
    ppm, pfm = tmp_path / "a.ppm", tmp_path / "a.pfm"
    k.write_ppm(img, ppm)
    k.write_pfm(img, pfm)
    assert ppm.read_bytes() == E.ppm_bytes(img)
    data = pfm.read_bytes()
    assert data == E.pfm_bytes(img)
    # read back: little-endian floats, bottom row first
    body = np.frombuffer(data[len(b"PF\n5 7\n-1\n"):], "<f4").reshape(7, 5, 3)
    assert np.array_equal(body[::-1], img)


def test_unwritable_image_is_an_error_with_the_path():
    img = np.zeros((2, 2, 3), np.float32)
    bad = os.path.join("/nonexistent-folder", "x.ppm")
    with pytest.raises(k.KodanukiError, match=r"x\.ppm:1:1: cannot write"):
        k.write_ppm(img, bad)


def test_scene_reads_grade_and_batch():
    head = 'import "standard.sdf";\nrender { width = 8; height = 8; batch = 16; }\n'
    s = k.Scene(head + "grade { exposure = 2; curve = clamped; }\nsphere(1);\n")
    _, rs = s.settings()
    assert (rs.batch, rs.exposure, rs.curve) == (16, 2.0, k.CURVE_CLAMPED)
    _, rs = k.Scene('import "standard.sdf";\nsphere(1);\n').settings()
    assert (rs.batch, rs.exposure, rs.curve) == (0, 1.0, k.CURVE_FILMIC)
    with pytest.raises(k.KodanukiError, match="filmic or clamped"):
        k.Scene(head + "grade { curve = x; }\nsphere(1);\n")
