"""The temporal member on the device (reconstruction/temporal), through the
C ABI:
- the device's temporal step (its LDS-staged neighbourhoods, the history's
  reprojection and blend) equals the reference's temporalFrame to the bit on
  the synthetic viewer sequence the fixtures hold, and the oracle's on
  another;
- a viewer on a rendered scene: with the camera still, each frame's history
  grows by a sample and its shown image converges toward a many-sample
  render of the same view; a step of the camera keeps most histories; a
  restart drops them; the denoised frame (the a-trous passes on the history)
  runs on the same buffers and keeps the image's mean."""
import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes
from oracle import temporal_ref as T
from tests import temporal_frames as TF
from tests.conftest import EMULATION
from tests.test_temporal import G, golden_frames

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)


def _lens(l):
    lens = k.Lens()
    lens.eye[:], lens.forward[:], lens.right[:], lens.up[:] = list(l[0]), list(l[1]), list(l[2]), list(l[3])
    lens.reach, lens.aspect = l[4], l[5]
    return lens


def _run_device(device, frames):
    before, out = None, []
    for fr in frames:
        px, img = device.temporal_step(_lens(fr["lens"]), _lens(fr["before_lens"]), fr["parts"], fr["record"], fr["gbuffer"],
                                       before=before, moved=fr["moved"], restarted=fr["restarted"], recounted=fr["recounted"])
        out.append((px, img))
        before = px
    return out


def test_device_step_equals_reference_vectors(device):
    for kf, (px, img) in enumerate(_run_device(device, golden_frames())):
        assert px.tobytes() == G[f"temporal{kf}_pixels"].tobytes(), kf
        assert img.tobytes() == G[f"temporal{kf}_image"].tobytes(), kf


@pytest.mark.parametrize("size", [(37, 21), (16, 16), (70, 45)])
def test_device_step_equals_the_oracle(device, size):
    W, H = size
    frames = TF.sequence(seed=W, width=W, height=H, frames=6)
    for kf, ((px, img), (opx, oimg)) in enumerate(zip(_run_device(device, frames), TF.run_oracle(frames))):
        assert px.tobytes() == opx.tobytes(), kf
        assert img.tobytes() == oimg.tobytes(), kf


def _viewer(device, W=48, H=32):
    s = k.Scene(scenes.stopping_table(width=W, height=H))
    cam, rs = s.settings()
    lower, size = s.region(6)
    device.upload(s)
    device.build_octree(lower, size, 6)
    return s, cam, rs


def test_still_viewer_converges(device):
    s, cam, rs = _viewer(device)
    W, H = 48, 32
    device.render_pass(W, H, 1024, first_sample=1 << 20, seed=rs.seed, surfaces=rs.surfaces)
    rgb, _, n = device.accumulation(W, H)
    truth = rgb / n[..., None]
    errs = []
    frames = 24 if EMULATION else 64
    for f in range(frames):
        device.render_pass(W, H, 1, first_sample=f, seed=rs.seed, surfaces=rs.surfaces, flags=k.PASS_RECORD)
        shown = device.temporal_frame(restarted=(f == 0))
        errs.append(float(np.mean(np.abs(shown - truth) / (truth + 0.05))))
    hist = device.temporal_history(W, H)
    note(f"[temporal] still viewer: error after 4 frames {errs[3]:.4f}, after {frames} {errs[-1]:.4f}; "
         f"lengths {np.percentile(hist['length'], [5, 50, 95])}")
    assert errs[-1] < 0.5 * errs[3]
    assert np.median(hist["length"]) == frames


def test_moving_viewer_keeps_histories_and_restart_drops_them(device):
    s, cam, rs = _viewer(device)
    W, H = 48, 32
    eye, target, up = list(cam.eye), list(cam.target), list(cam.up)
    for f in range(4):
        device.render_pass(W, H, 1, first_sample=f, seed=rs.seed, surfaces=rs.surfaces, flags=k.PASS_RECORD)
        device.temporal_frame(restarted=(f == 0))
    eye2 = [eye[0] + 0.02, eye[1], eye[2]]
    device.set_camera(eye2, target, up, cam.fov)
    device.render_pass(W, H, 1, first_sample=4, seed=rs.seed, surfaces=rs.surfaces, flags=k.PASS_RECORD)
    (shown, denoised), st = device.temporal_frame(moved=True, denoise=True, stats=True)
    hist = device.temporal_history(W, H)
    kept = np.mean(hist["length"] > 1)
    note(f"[temporal] a step of 2 cm keeps {kept:.3f} of the histories; gbuffer {st.gbuffer_ms:.2f} ms, temporal "
         f"{st.temporal_ms:.3f} ms, filter {st.filter_ms:.2f} ms; mean shown {shown.mean():.4f} denoised {denoised.mean():.4f}")
    assert kept > 0.8
    assert abs(denoised.mean() - shown.mean()) < 0.05 * shown.mean()
    device.render_pass(W, H, 1, first_sample=5, seed=rs.seed, surfaces=rs.surfaces, flags=k.PASS_RECORD)
    device.temporal_frame(restarted=True)
    assert device.temporal_history(W, H)["length"].max() <= 1
    device.set_camera(eye, target, up, cam.fov)
