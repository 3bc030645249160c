"""The oracle pinned to the reference's own functions.

tests/golden/ref_vectors.npz holds inputs and outputs of the reference's
functions (compiled from its sources into oracle/_ref by `make -C oracle
ref`; written by tests/golden/make_ref_vectors.py). The oracle's
restatements must reproduce them: the Sobol member (sobol.cpp), the hashed
numbers (random.cpp), the wavelength cuts (colour.cpp:134-257), colourOf and
spectrumOf, the GGX escape tables (energy.cpp:100-127), effortFor,
splitCeiling, the estimate fold and splitSpread (core/stopping.cpp), and the
grade, byte encoding, PPM and PFM (shared/image.cpp). Where the built
reference library is at hand, fresh random keys are checked against it too.

The unscrambled Sobol points are checked against scipy's own Sobol generator
(an independent implementation of Joe and Kuo's sequence, which draws points
in Gray-code order), and the scrambled ones against the reference test's net
and evenness checks (target/tests/sampler.cpp:378-435).
"""
import ctypes as C
import os
import tempfile

import numpy as np
import pytest

import kodanuki_amd as k
from oracle import estimate_ref as E
from oracle import sobol_ref as S
from oracle import stopping_ref as T
from tests import oracle_py as O

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
G = np.load(os.path.join(HERE, "golden", "ref_vectors.npz"))
REF = os.path.join(ROOT, "oracle", "_ref", "libkdn_ref.so")


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def test_sobol_matches_reference_vectors():
    got = S.sobol_number(G["sobol_stream"], G["sobol_index"], G["sobol_coordinate"])
    assert np.array_equal(got, G["sobol_value"])


def test_sobol_c_oracle_matches_numpy():
    L = O.lib()
    got = np.array([L.or_sobol_number(int(s), int(i), int(c))
                    for s, i, c in zip(G["sobol_stream"][:1024], G["sobol_index"][:1024], G["sobol_coordinate"][:1024])])
    assert np.array_equal(got, G["sobol_value"][:1024])


def test_sobol_unscrambled_matches_scipy():
    qmc = pytest.importorskip("scipy.stats.qmc")
    points = qmc.Sobol(d=28, scramble=False).random_base2(12)
    n = np.arange(len(points), dtype=np.uint64)
    gray = n ^ (n >> np.uint64(1))
    mine = np.stack([S.unscrambled(d, gray) for d in range(28)], axis=1)
    assert np.array_equal(mine, points)


@pytest.mark.parametrize("stream", [0x0123456789ABCDEF, 0xFEDCBA9876543210, 0x5555AAAA33331111])
def test_sobol_nets_and_evenness(stream):
    """Every power-of-two prefix of the first pair of the shared and of the
    own coordinates is a (0, k, 2)-net, and of every coordinate holds one
    point per interval (the reference's sobolNet and sobolEven)."""
    for k_ in range(1, 9):
        count = 1 << k_
        idx = np.arange(count, dtype=np.uint64)
        st = np.full(count, stream, dtype=np.uint64)
        for dim in (0, S.SHARED_COORDINATES):
            x = S.sobol_number(st, idx, np.full(count, dim, np.uint64))
            y = S.sobol_number(st, idx, np.full(count, dim + 1, np.uint64))
            for a in range(k_ + 1):
                cell = (np.floor(y * (1 << (k_ - a))).astype(np.int64) << a) + np.floor(x * (1 << a)).astype(np.int64)
                assert np.array_equal(np.sort(cell), np.arange(count)), (dim, k_, a)
        for dim in range(28):
            v = S.sobol_number(st, idx, np.full(count, dim, np.uint64))
            assert np.array_equal(np.sort(np.floor(v * count).astype(np.int64)), np.arange(count)), (dim, k_)


def test_sobol_no_two_samples_on_one_point():
    idx = np.arange(1 << 16, dtype=np.uint64)
    v = S.sobol_number(np.full(len(idx), 0x9E3779B97F4A7C15, np.uint64), idx, np.zeros(len(idx), np.uint64))
    assert len(np.unique(v)) == len(idx)


def test_hashed_numbers_match_reference_vectors():
    L = O.lib()
    got = np.array([L.or_hashed_number(int(s), int(i), int(c))
                    for s, i, c in zip(G["hashed_stream"], G["hashed_index"], G["hashed_coordinate"])])
    assert np.array_equal(got, G["hashed_value"])


def test_wavelength_cuts_match_reference_vectors():
    L = O.lib()
    L.or_sampled_wavelengths.argtypes = [C.c_double, C.c_void_p, C.c_void_p]
    nm = np.zeros(4)
    dens = np.zeros(4)
    for u, rnm, rd in zip(G["wave_u"], G["wave_nm"], G["wave_density"]):
        L.or_sampled_wavelengths(u, ptr(nm), ptr(dens))
        assert np.array_equal(nm.astype(np.float32), rnm)
        assert np.array_equal(dens.astype(np.float32), rd)


def test_colour_and_spectrum_match_reference_vectors():
    L = O.lib()
    L.or_colour_of.argtypes = [C.c_void_p] * 4
    L.or_spectrum_of.argtypes = [C.c_void_p] * 3
    rgb = np.zeros(3)
    back = np.zeros(4)
    for k_ in range(len(G["colour_rgb"])):
        nm = np.ascontiguousarray(G["colour_nm"][k_])
        L.or_colour_of(ptr(nm), ptr(np.ascontiguousarray(G["colour_density"][k_])), ptr(np.ascontiguousarray(G["colour_spectrum"][k_])),
                       ptr(rgb))
        assert np.array_equal(rgb, G["colour_rgb"][k_])
        L.or_spectrum_of(ptr(np.ascontiguousarray(G["spectrum_rgb"][k_])), ptr(nm), ptr(back))
        # the reference keeps its basis curves in float: within float rounding
        assert np.abs(back - G["spectrum_value"][k_]).max() < 1e-7


def test_escape_tables_match_reference_vectors():
    L = O.lib()
    for w, table, avg in zip(G["escape_width"], G["escape_table"], G["escape_average"]):
        t = np.zeros(32)
        a = np.zeros(1)
        L.or_escape_of(w, ptr(t), ptr(a))
        assert np.array_equal(t, table) and a[0] == avg


def test_effort_and_split_ceiling_match_reference_vectors():
    for t, (samples, batch) in zip(G["effort_tolerance"], G["effort"]):
        assert T.effort_for(float(t)) == (samples, batch)
        assert k.stopping_effort(float(t)) == (samples, batch)
    for b, c in zip(G["ceiling_batch"], G["ceiling"]):
        assert T.split_ceiling(int(b)) == c


def _packed(e):
    return np.array(list(e["total"]) + [e["taken"]] + list(e["batch"]) + [e["batches"]] + list(e["means"]) + list(e["squares"]) +
                    [e["shared"], e["shared_means"], e["shared_squares"], e["own_squares"]])


def test_fold_and_split_spread_match_reference_vectors():
    for colours, state, spread in zip(G["fold_colours"], G["fold_state"], G["fold_spread"]):
        e = E.fold_pixel(E.empty(), colours, 8)
        assert np.array_equal(_packed(e), state)
        rec = E.as_record(e, k.ESTIMATE_DTYPE)
        assert T.split_spread(rec, 8, 1) == (spread[0], spread[1])


def test_grade_matches_reference_vectors():
    for a, b, e, c, s in zip(G["steps_low"], G["steps_high"], G["steps_exposure"], G["steps_curve"], G["steps"]):
        assert T.steps_between(float(a), float(b), float(e), int(c)) == s
    enc = k.encode_image(G["steps_low"].astype(np.float32).reshape(1, -1, 1).repeat(3, axis=2)[:, :, :], 1.0, 0)
    # the library encodes float pixels: compare where the float is the double
    exact = G["steps_low"].astype(np.float32).astype(np.float64) == G["steps_low"]
    ones = (G["steps_exposure"] == 1.0) & (G["steps_curve"] == 0) & exact
    assert np.array_equal(enc[0, ones, 0], G["byte"][ones])


def test_ppm_and_pfm_match_reference_files():
    with tempfile.TemporaryDirectory() as d:
        k.write_ppm(G["image"], os.path.join(d, "a.ppm"), exposure=1.5, curve=k.CURVE_FILMIC)
        k.write_pfm(G["image"], os.path.join(d, "a.pfm"))
        with open(os.path.join(d, "a.ppm"), "rb") as f:
            assert f.read() == G["ppm"].tobytes()
        with open(os.path.join(d, "a.pfm"), "rb") as f:
            assert f.read() == G["pfm"].tobytes()


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref is not built (make -C oracle ref, needs the reference tree)")
def test_fresh_keys_match_built_reference():
    L = C.CDLL(REF)
    L.kr_sobol_number.restype = C.c_double
    L.kr_sobol_number.argtypes = [C.c_uint64] * 3
    rng = np.random.default_rng()
    streams = rng.integers(0, 2 ** 63, 2000, dtype=np.int64).astype(np.uint64) * np.uint64(2)
    index = rng.integers(0, 1 << 32, 2000, dtype=np.int64).astype(np.uint64)
    coord = rng.integers(0, 28, 2000).astype(np.uint64)
    ref = np.array([L.kr_sobol_number(int(s), int(i), int(c)) for s, i, c in zip(streams, index, coord)])
    assert np.array_equal(S.sobol_number(streams, index, coord), ref)
