"""The largest configuration's sizes on the device: the production scene
grown to 4096 primitives under an octree of depth 10, paths of up to 32
surfaces, a 3840x2160 image. Checked by what does not depend on size: hits
within 1e-4 relative of the FP64 sphere trace of the whole scene (the
octree twin is not built at this size), a pass of two samples equal to the
bit to two progressive passes of one, and the finishing launch equal to a
launch per bounce on a window of the image."""
import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes
from tests import oracle_py as O
from tests import raysets
from tests.conftest import EMULATION

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(bool(EMULATION), reason="too large for the CPU build")]

W, H, SURFACES, DEPTH = 3840, 2160, 32, 10


def note(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def large(device):
    s = k.Scene(scenes.production(3, count=4096, width=W, height=H, samples=256, surfaces=SURFACES, glass=False))
    lower, size = s.region(DEPTH)
    device.upload(s)
    st = device.build_octree(lower, size, DEPTH)
    note(f"[large] 4096 primitives, octree depth {DEPTH}: {st.entries} entries, {st.pool_steps} pool steps, "
         f"build {st.build_ms:.1f} ms")
    return s, lower, size, st


def test_hits_against_fp64_sphere_trace(device, large):
    s, lower, size, _ = large
    A = O.SceneArrays(s)
    rays = np.concatenate([raysets.camera_rays(s, 128, 5), raysets.random_rays(A, lower, size, 128, 6, inside_share=0.0)])
    hits = device.extend(rays)
    ref, d64 = O.sphere_trace64(A, lower, size, rays)
    agree = np.mean((hits["kind"] != k.KDN_HIT_MISSED) == (ref["kind"] != k.KDN_HIT_MISSED))
    both = (hits["kind"] == k.KDN_HIT_SURFACE) & (ref["kind"] == k.KDN_HIT_SURFACE)
    rel = np.abs(hits["distance"][both].astype(np.float64) - d64[both]) / np.maximum(d64[both], 0.05)
    facing = np.abs(np.sum(rays["direction"][both].astype(np.float64) * ref["normal"][both].astype(np.float64), axis=1))
    note(f"[large] hit/miss agreement {agree:.4f}, {both.sum()} hits, max rel {rel.max():.2e}, "
         f"leaf agreement {np.mean(hits['leaf'][both] == ref['leaf'][both]):.4f}")
    assert agree > 0.99 and both.sum() > 50
    assert np.all(rel[facing >= 0.01] < 1e-4)
    assert np.mean(hits["leaf"][both] == ref["leaf"][both]) > 0.98


def test_progressive_passes_equal_one_pass(device, large):
    s, lower, size, _ = large
    seed = s.settings()[1].seed
    st = device.render_pass(W, H, 2, seed=seed, surfaces=SURFACES)
    a, sa, na = device.accumulation(W, H)
    device.render_pass(W, H, 1, seed=seed, surfaces=SURFACES)
    device.render_pass(W, H, 1, first_sample=1, seed=seed, surfaces=SURFACES, keep=True)
    b, sb, nb = device.accumulation(W, H)
    note(f"[large] {W}x{H} x 2 spp: {st.ms:.0f} ms, {st.segments} segments, {st.shadow_rays} shadow rays, "
         f"mean {(a / na[..., None]).mean(axis=(0, 1))}")
    # (linear sRGB of a spectral sample may be negative: the colour
    # matrix has negative terms)
    assert np.all(na == 2) and np.all(np.isfinite(a)) and np.all(np.isfinite(sa))
    assert st.segments > W * H * 2
    assert np.array_equal(na, nb)
    assert np.allclose(a, b, rtol=0, atol=1e-9) and np.allclose(sa, sb, rtol=0, atol=1e-9)


def test_finishing_launch_equals_a_launch_per_bounce(device, large):
    s, lower, size, _ = large
    seed = s.settings()[1].seed
    w, h = 480, 270
    a_st = device.render_pass(w, h, 4, seed=seed, surfaces=SURFACES)
    a, sa, _ = device.accumulation(w, h)
    b_st = device.render_pass(w, h, 4, seed=seed, surfaces=SURFACES, flags=k.PASS_NO_FINISH)
    b, sb, _ = device.accumulation(w, h)
    note(f"[large] {w}x{h} x 4: segments {a_st.segments} / {b_st.segments}, bounces {b_st.bounces}")
    assert a.tobytes() == b.tobytes() and sa.tobytes() == sb.tobytes()
    assert (a_st.segments, a_st.shadow_rays) == (b_st.segments, b_st.shadow_rays)
