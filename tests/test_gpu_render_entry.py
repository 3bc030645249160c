"""The render entry point and the plugin-order C++ host on the device.

The C++ host (tests/host/render_host.cpp) loads the HIP library as a plugin's
shared object and draws a still frame in the order the core calls a backend's
hooks; its image is the Python path's to the byte. The TOML-driven entry point
(kodanuki_amd.render) draws the frame its config asks for: with fixed_count
and the mean reconstruction it is the same frame; with the Sobol member and
the a-trous filter it is the library's Sobol pass and denoised accumulation.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import render as R
from kodanuki_amd import scenes

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(bool(os.environ.get("KDN_EMULATION")), reason="the device's library")]

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BASE = ('"backend/cpu" = false\n"backend/hip" = true\n"geometry/tree" = false\n"geometry/hip" = true\n'
        '"execution/packets" = false\n"execution/hip" = true\n"sampler/highlight_guide" = false\n')


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def _mean_frame(dev, scene_path, w, h, spp, depth, flags=0):
    scene = k.Scene.from_file(scene_path)
    cam, rs = scene.settings()
    lower, size = scene.region(depth)
    dev.upload(scene)
    dev.build_octree(lower, size, depth)
    dev.render_pass(w, h, spp, seed=rs.seed, surfaces=rs.surfaces, flags=flags | k.PASS_ESTIMATES, batch=spp)
    est = dev.estimates(w, h)
    return (est["total"] / np.maximum(est["taken"], 1)[..., None]).astype(np.float32), rs


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_in_plugin_order_matches_the_python_path(device, tmp_path):
    scene = _write(tmp_path, "tiny.sdf", scenes.tiny(glass=False))
    exe = str(tmp_path / "render_host")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "host",
                    "render_host.cpp"), "-o", exe, "-ldl"], check=True)
    print("[host] 64x48, 4 spp, depth 5", flush=True)
    host = subprocess.run([exe, k.LIB_PATH, scene, str(tmp_path / "host"), "64", "48", "4", "5"], capture_output=True, text=True,
                          timeout=120)
    assert host.returncode == 0, host.stderr[-2000:]
    shown, rs = _mean_frame(device, scene, 64, 48, 4, 5)
    k.write_ppm(shown, str(tmp_path / "py.ppm"), exposure=rs.exposure, curve=rs.curve)
    assert (tmp_path / "host.ppm").read_bytes() == (tmp_path / "py.ppm").read_bytes()


def test_entry_point_draws_what_its_config_asks(device, tmp_path):
    scene = _write(tmp_path, "tiny.sdf", scenes.tiny(glass=False))
    fixed = _write(tmp_path, "fixed.toml", '[settings]\nwidth = 64\nheight = 48\nsamples = 8\n[plugins]\n' + BASE +
                   '"stopping/confidence_interval" = false\n"stopping/fixed_count" = true\n"sampler/sobol" = false\n')
    shown, written, plan = R.render(scene, fixed, out_dir=str(tmp_path / "out"), depth=5, dev=device)
    assert [os.path.basename(p) for p in written] == ["tiny.ppm", "tiny.pfm"]
    ref, _ = _mean_frame(device, scene, 64, 48, 8, 5)
    assert shown.tobytes() == ref.tobytes()
    sobol = _write(tmp_path, "sobol.toml", '[settings]\nwidth = 64\nheight = 48\nsamples = 8\n[plugins]\n' + BASE +
                   '"stopping/confidence_interval" = false\n"stopping/fixed_count" = true\n"sampler/sobol" = true\n'
                   '"reconstruction/atrous" = true\n')
    shown, _, plan = R.render(scene, sobol, out_dir=str(tmp_path / "out2"), depth=5, dev=device)
    assert plan["sobol"] and plan["denoise"]
    _mean_frame(device, scene, 64, 48, 8, 5, flags=k.PASS_SOBOL)
    ref = device.denoise_accumulation(64, 48)
    assert shown.tobytes() == ref.tobytes()
