"""Config 4 on the device: the stress scene (the production scene grown to
4096 primitives with a 64 x 64 canopy of repeated bubbles, each a thin
glass shell around a nested dielectric core; kodanuki_amd.scenes.stress)
under an octree of depth 10, and the viewer sequence drawn on it (one
sample a frame, a seeded orbit with jitter of the eye, the temporal member
and the a-trous filter on every frame).

Glass is held off the device, so the renders take the scene's matte
variant (same geometry); the hits are checked against the FP64 sphere trace
of the whole scene (the octree twin is not built at this size) for camera
rays and rays from outside every solid.
"""
import math

import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes
from tests import oracle_py as O
from tests import raysets
from tests.conftest import EMULATION

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(bool(EMULATION), reason="too large for the CPU build")]

DEPTH = 10


def note(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def stress(device):
    s = k.Scene(scenes.stress(0, glass=False))
    lower, size = s.region(DEPTH)
    note(f"[stress] {s.info.leaves} leaves, {s.info.objects} objects, octree depth {DEPTH}")
    device.upload(s)
    st = device.build_octree(lower, size, DEPTH)
    note(f"[stress] {st.entries} entries, {st.pool_steps} pool steps, build {st.build_ms:.1f} ms")
    return s, lower, size, st


def test_hits_against_fp64_sphere_trace(device, stress):
    s, lower, size, _ = stress
    A = O.SceneArrays(s)
    rays = np.concatenate([raysets.camera_rays(s, 64, 5), raysets.random_rays(A, lower, size, 64, 6, inside_share=0.0)])
    hits = device.extend(rays)
    ref, d64 = O.sphere_trace64(A, lower, size, rays)
    agree = np.mean((hits["kind"] != k.KDN_HIT_MISSED) == (ref["kind"] != k.KDN_HIT_MISSED))
    both = (hits["kind"] == k.KDN_HIT_SURFACE) & (ref["kind"] == k.KDN_HIT_SURFACE)
    rel = np.abs(hits["distance"][both].astype(np.float64) - d64[both]) / np.maximum(d64[both], 0.05)
    facing = np.abs(np.sum(rays["direction"][both].astype(np.float64) * ref["normal"][both].astype(np.float64), axis=1))
    note(f"[stress] hit/miss agreement {agree:.4f}, {both.sum()} hits, max rel {rel.max():.2e}, "
         f"leaf agreement {np.mean(hits['leaf'][both] == ref['leaf'][both]):.4f}")
    assert agree > 0.98 and both.sum() > 40
    assert np.all(rel[facing >= 0.01] < 1e-4)
    assert np.mean(hits["leaf"][both] == ref["leaf"][both]) > 0.97


def test_progressive_passes_equal_one_pass(device, stress):
    s, _, _, _ = stress
    rs = s.settings()[1]
    W, H = 480, 270
    st = device.render_pass(W, H, 2, seed=rs.seed, surfaces=rs.surfaces)
    a, sa, na = device.accumulation(W, H)
    device.render_pass(W, H, 1, seed=rs.seed, surfaces=rs.surfaces)
    device.render_pass(W, H, 1, first_sample=1, seed=rs.seed, surfaces=rs.surfaces, keep=True)
    b, sb, nb = device.accumulation(W, H)
    note(f"[stress] {W}x{H} x 2 spp, {rs.surfaces} surfaces: {st.ms:.1f} ms, {st.segments} segments")
    assert np.all(na == 2) and np.all(np.isfinite(a))
    assert np.array_equal(na, nb) and np.allclose(a, b, rtol=0, atol=1e-9) and np.allclose(sa, sb, rtol=0, atol=1e-9)


def viewer_sequence(device, scene, width, height, frames, seed=4):
    """The config-4 viewer: a frame of one sample each, the eye orbiting
    the target by 0.2 degrees a frame with a seeded jitter of a centimetre,
    the temporal member then the a-trous filter on every frame. Returns the
    frames' (shown, denoised) means, the last frame's history and the
    seconds of the sequence."""
    import time
    cam, rs = scene.settings()
    eye, target, up = np.array(cam.eye[:], np.float64), np.array(cam.target[:], np.float64), list(cam.up)
    rng = np.random.default_rng(seed)
    arm = eye - target
    means = []
    start = time.perf_counter()
    for f in range(frames):
        a = math.radians(0.2) * f
        rot = np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])
        e = target + rot @ arm + rng.normal(0, 0.01, 3)
        device.set_camera(list(e), list(target), up, cam.fov)
        device.render_pass(width, height, 1, first_sample=f, seed=rs.seed, surfaces=rs.surfaces, flags=k.PASS_RECORD)
        shown, denoised = device.temporal_frame(moved=f > 0, restarted=f == 0, denoise=True)
        means.append((float(shown.mean()), float(denoised.mean())))
    seconds = time.perf_counter() - start
    device.set_camera(list(eye), list(target), up, cam.fov)
    return means, device.temporal_history(width, height), seconds


def test_viewer_sequence(device, stress):
    s, _, _, _ = stress
    W, H, F = 480, 270, 12
    means, hist, seconds = viewer_sequence(device, s, W, H, F)
    note(f"[stress] viewer {W}x{H} x {F} frames in {seconds:.2f} s; history lengths {np.percentile(hist['length'], [5, 50, 95])}; "
         f"last frame mean shown {means[-1][0]:.4f} denoised {means[-1][1]:.4f}")
    assert all(np.isfinite(m).all() for m in means)
    # a slow orbit keeps most histories: they grow toward the frame count
    assert np.median(hist["length"]) >= F / 2
    assert abs(means[-1][1] - means[-1][0]) < 0.05 * abs(means[-1][0])
