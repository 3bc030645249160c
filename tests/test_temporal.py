"""The temporal oracle (oracle/temporal_ref.py) against the reference.

- To the bit against the reference's own temporalFrame on a synthetic
  viewer sequence (restart, still frames, sideways steps, an orbit step and
  a recounted frame): fixtures written from oracle/_ref
  (tests/golden/make_ref_vectors.py), and fresh sequences against the built
  reference where it is at hand.
- The reference test's behaviours (target/tests/reconstruction.cpp): a
  still camera's noisy samples converge to their mean and its history grows
  past the moving cap; a sideways step carries the history to where its
  points show; a pixel showing another leaf, or whose distance jumps, drops
  its history; a restart and a new size drop every history; a pixel holds
  the G-buffer's albedo, the sky none.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from oracle import temporal_ref as T
from tests import temporal_frames as TF

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libkdn_ref.so")
G = np.load(os.path.join(HERE, "golden", "ref_vectors.npz"))


def golden_frames():
    frames = []
    before = None
    for k in range(8):
        lv = G[f"temporal{k}_lens"]
        lens = (tuple(lv[0:3]), tuple(lv[3:6]), tuple(lv[6:9]), tuple(lv[9:12]), float(lv[12]), float(lv[13]))
        moved, restarted, recounted = (bool(v) for v in G[f"temporal{k}_flags"])
        parts = G[f"temporal{k}_parts"]
        H, W = parts.shape[:2]
        frames.append(dict(width=W, height=H, lens=lens, before_lens=before or lens, moved=moved, restarted=restarted,
                           recounted=recounted, parts=parts, record=G[f"temporal{k}_record"], gbuffer=G[f"temporal{k}_gbuffer"]))
        before = lens
    return frames


def test_oracle_matches_reference_vectors():
    for k, (px, img) in enumerate(TF.run_oracle(golden_frames())):
        assert px.tobytes() == G[f"temporal{k}_pixels"].tobytes(), k
        assert img.tobytes() == G[f"temporal{k}_image"].tobytes(), k


@pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref is not built (make -C oracle ref, needs the reference tree)")
@pytest.mark.parametrize("seed", [11, 12])
def test_oracle_matches_built_reference(seed):
    from tests.golden import make_ref_vectors as M
    L = M.ref()
    L.kr_temporal_frame.argtypes = [C.c_int] * 2 + [C.c_void_p] * 2 + [C.c_int] * 3 + [C.c_void_p] * 3 + [C.c_int] * 3 + [C.c_void_p] * 3
    frames = TF.sequence(seed=seed, width=20, height=14, frames=7)
    before, bw, bh, bd = np.zeros((1, 1), T.PIXEL_DTYPE), 0, 0, 0
    for fr, (mine, img) in zip(frames, TF.run_oracle(frames)):
        W, H = fr["width"], fr["height"]
        px = np.zeros((H, W), T.PIXEL_DTYPE)
        im = np.zeros((H, W, 3), np.float32)
        L.kr_temporal_frame(W, H, C.byref(M.lens_struct(fr["lens"], W, H)), C.byref(M.lens_struct(fr["before_lens"], W, H)),
                            int(fr["moved"]), int(fr["restarted"]), int(fr["recounted"]), M.ptr(fr["parts"]), M.ptr(fr["record"]),
                            M.ptr(fr["gbuffer"]), bw, bh, bd, M.ptr(before), M.ptr(px), M.ptr(im))
        assert px.tobytes() == mine.tobytes() and im.tobytes() == img.tobytes()
        before, bw, bh, bd = px, W, H, 1


def test_still_camera_converges():
    """Noisy samples on a still camera converge to their mean (the error
    late under 0.3 of the error early) and the history grows by a sample a
    frame past the moving cap."""
    rng = np.random.default_rng(4)
    W, H = 12, 10
    lens = TF.lens((0, -4, 2.5), (0, 1, 0), width=W, height=H)
    base = TF.frame(lens, lens, W, H, np.random.default_rng(9), noise=0.0)
    truth = base["parts"]
    state, early, late = None, None, None
    for f in range(120):
        fr = dict(base, restarted=(f == 0), record=base["gbuffer"],
                  parts=truth * (1 + 0.5 * rng.standard_normal(truth.shape[:2]))[..., None])
        px, img = T.temporal_frame(state, fr)
        state = dict(width=W, height=H, demodulated=True, pixels=px)
        err = float(np.mean(np.abs(img - truth)))
        early = err if f == 3 else early
        late = err
    assert late < 0.3 * early
    assert np.all(state["pixels"]["length"] == 120)


def test_sideways_step_carries_history():
    W, H = 32, 24
    rng = np.random.default_rng(5)
    a = TF.lens((0, -4, 2.5), (0, 1, 0), width=W, height=H)
    b = TF.lens((0.05, -4, 2.5), (0.05, 1, 0), width=W, height=H)
    first = TF.frame(a, a, W, H, rng, restarted=True)
    px0, _ = T.temporal_frame(None, first)
    second = TF.frame(b, a, W, H, rng, moved=True)
    px1, _ = T.temporal_frame(dict(width=W, height=H, demodulated=True, pixels=px0), second)
    # most floor and wall pixels found the pixels their points showed before
    surfaced = px1["leaf"] != T.NO_LEAF
    assert np.mean(px1["length"][surfaced] > 1) > 0.8
    # the sky carries its history too (the sky matches the sky alone)
    assert np.all(px1["length"][~surfaced][1:] >= 1)


def test_other_leaf_or_distance_jump_drops_history():
    W, H = 24, 16
    rng = np.random.default_rng(6)
    a = TF.lens((0, -4, 2.5), (0, 1, 0), width=W, height=H)
    b = TF.lens((0.02, -4, 2.5), (0.02, 1, 0), width=W, height=H)
    first = TF.frame(a, a, W, H, rng, restarted=True)
    px0, _ = T.temporal_frame(None, first)
    second = TF.frame(b, a, W, H, rng, moved=True)
    second["record"] = second["gbuffer"].copy()
    g = second["gbuffer"]
    # a pixel now showing another leaf, and one whose distance jumps
    g[8, 5]["leaf"] = 9
    second["record"][8, 5]["leaf"] = 9
    g[8, 15]["distance"] = g[8, 15]["distance"] * 0.5
    second["record"][8, 15]["distance"] = g[8, 15]["distance"]
    px1, _ = T.temporal_frame(dict(width=W, height=H, demodulated=True, pixels=px0), second)
    assert px1[8, 5]["length"] == 1 and px1[8, 15]["length"] == 1
    assert px1[8, 10]["length"] == 2


def test_restart_and_resize_drop_every_history():
    W, H = 16, 12
    rng = np.random.default_rng(7)
    a = TF.lens((0, -4, 2.5), (0, 1, 0), width=W, height=H)
    state = None
    for f in range(3):
        px, _ = T.temporal_frame(state, TF.frame(a, a, W, H, rng, restarted=(f == 0)))
        state = dict(width=W, height=H, demodulated=True, pixels=px)
    assert px["length"].max() == 3
    px, _ = T.temporal_frame(state, TF.frame(a, a, W, H, rng, restarted=True))
    assert px["length"].max() == 1
    small = TF.lens((0, -4, 2.5), (0, 1, 0), width=8, height=6)
    px, _ = T.temporal_frame(state, TF.frame(small, small, 8, 6, rng))
    assert px["length"].max() == 1


def test_pixels_hold_the_gbuffer_albedo():
    W, H = 16, 12
    rng = np.random.default_rng(8)
    a = TF.lens((0, -4, 2.5), (0, 1, 0), width=W, height=H)
    fr = TF.frame(a, a, W, H, rng, restarted=True)
    px, img = T.temporal_frame(None, fr)
    assert np.array_equal(px["albedo"], fr["gbuffer"]["albedo"].astype(np.float64))
    sky = fr["gbuffer"]["leaf"] == T.NO_LEAF
    assert np.all(px["albedo"][sky] == 0)
    # the shown light is the lighting times the albedo (1 where below 1e-3)
    d = np.where(px["albedo"] < 1e-3, 1.0, px["albedo"])
    assert np.array_equal(img, (px["colour"] * d).astype(np.float32))
