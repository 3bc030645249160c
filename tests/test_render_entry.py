"""The render entry point and the plugin-order host of the C ABI.

- kodanuki_amd.render checks a TOML render config as the reference's
  workflow.py does (target/workflow.py:48-83: known sections and keys, true
  or false plugins, exactly one geometry, execution, backend and stopping)
  with `hip` as a geometry, execution and backend member, and maps it onto
  the library's pass (stopping, Sobol, plain guide, a-trous, sink/file);
- tests/host/render_host.cpp, a C++ program that loads the library as a
  plugin's shared object and calls it in the order the core calls a
  backend's hooks (open, build the world, a round with estimates, read the
  estimates, write PPM and PFM, close), builds against include/kodanuki.h
  and, run on the bounds-checked CPU build of the library, writes the same
  image bytes as the Python path through the same library.
"""
import os
import shutil
import subprocess
import sys

import pytest

from kodanuki_amd import render as R
from kodanuki_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CONFIGS = os.path.join(ROOT, "kodanuki_amd", "configs")
HOST_SRC = os.path.join(ROOT, "tests", "host", "render_host.cpp")
EMU = os.path.join(ROOT, "tests", "cpu_build")


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_committed_configs_check_and_plan():
    plan = R.gpu_plan(R.config_of(os.path.join(CONFIGS, "render_hip.toml")))
    assert plan["stopping"] == "confidence_interval" and plan["sobol"] and plan["file"] and not plan["denoise"]
    plan = R.gpu_plan(R.config_of(os.path.join(CONFIGS, "render_hip_fast.toml")))
    assert plan["stopping"] == "fixed_count" and plan["denoise"] and plan["settings"]["samples"] == 16


@pytest.mark.parametrize("text,message", [
    ('[settings]\nwidth = 64\n[plugins]\n"backend/hip" = true\n"backend/cpu" = false\n', "geometry/hip and execution/hip"),
    ('[settings]\nwidht = 64\n', "unknown key widht"),
    ('[look]\n', "unknown key look"),
    ('[plugins]\n"backend/cpu" = true\n', "backend/hip"),
    ('[plugins]\n"backend/hip" = true\n', "exactly one backend"),
    ('[plugins]\n"transport/paths" = false\n', "cannot be left out"),
    ('[plugins]\n"sampler/sobol" = 1\n', "true or false"),
    ('[plugins]\n"stopping/fixed_count" = true\n', "exactly one stopping"),
])
def test_configs_are_refused_as_the_workflow_refuses_them(tmp_path, text, message):
    path = _write(tmp_path, "c.toml", text)
    with pytest.raises(R.ConfigError, match=message):
        R.gpu_plan(R.config_of(path))


def test_gpu_path_refuses_what_it_cannot_draw(tmp_path):
    base = ('[plugins]\n"backend/cpu" = false\n"backend/hip" = true\n"geometry/tree" = false\n"geometry/hip" = true\n'
            '"execution/packets" = false\n"execution/hip" = true\n')
    assert R.gpu_plan(R.config_of(_write(tmp_path, "ok.toml", base)))["stopping"] == "confidence_interval"
    with pytest.raises(R.ConfigError, match="split"):
        R.gpu_plan(R.config_of(_write(tmp_path, "s.toml", "[settings]\nsplit = 4\n" + base)))
    with pytest.raises(R.ConfigError, match="viewer or a sequence"):
        R.gpu_plan(R.config_of(_write(tmp_path, "t.toml", base + '"reconstruction/temporal" = true\n')))


def _asan():
    try:
        out = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return out if os.path.isabs(out) and os.path.exists(out) else None


def build_host(out_dir):
    exe = os.path.join(str(out_dir), "render_host")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), HOST_SRC, "-o", exe, "-ldl"], check=True)
    return exe


# the Python path through the same library, in its own process (the CPU
# build is bound in place of the HIP library)
_PY_FRAME = r"""
import ctypes, sys
import numpy as np
sys.path.insert(0, {root!r})
import kodanuki_amd as k
k._bind(ctypes.CDLL({lib!r}))
scene = k.Scene.from_file({scene!r})
cam, rs = scene.settings()
lower, size = scene.region({depth})
dev = k.Device(0)
dev.upload(scene)
dev.build_octree(lower, size, {depth})
dev.render_pass({w}, {h}, {spp}, seed=rs.seed, surfaces=rs.surfaces, flags=k.PASS_ESTIMATES, batch={spp})
est = dev.estimates({w}, {h})
shown = (est["total"] / np.maximum(est["taken"], 1)[..., None]).astype(np.float32)
k.write_ppm(shown, {out!r} + ".ppm", exposure=rs.exposure, curve=rs.curve)
from kodanuki_amd import render as R
shown2, written, plan = R.render({scene!r}, {config!r}, out_dir={outdir!r}, depth={depth}, dev=dev)
assert plan["stopping"] == "fixed_count" and not plan["sobol"]
dev.close()
print("python frame written")
"""


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_host_in_plugin_order_matches_the_python_path_on_the_cpu_build(tmp_path):
    asan = _asan()
    if asan is None:
        pytest.skip("no AddressSanitizer runtime")
    subprocess.run(["make", "-s", "-C", EMU, "-j4", "libkodanuki_emu.so"], check=True, timeout=900)
    lib = os.path.join(EMU, "libkodanuki_emu.so")
    scene = _write(tmp_path, "tiny.sdf", scenes.tiny(glass=False))
    config = _write(tmp_path, "fixed.toml",
                    '[settings]\nwidth = 32\nheight = 24\nsamples = 2\n[plugins]\n"backend/cpu" = false\n"backend/hip" = true\n'
                    '"geometry/tree" = false\n"geometry/hip" = true\n"execution/packets" = false\n"execution/hip" = true\n'
                    '"stopping/confidence_interval" = false\n"stopping/fixed_count" = true\n"sampler/sobol" = false\n'
                    '"sampler/highlight_guide" = false\n')
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    env.pop("KDN_EMULATION", None)
    exe = build_host(tmp_path)
    host = subprocess.run([exe, lib, scene, str(tmp_path / "host"), "32", "24", "2", "5"], env=env, capture_output=True,
                          text=True, timeout=600)
    assert host.returncode == 0, host.stderr[-3000:]
    code = _PY_FRAME.format(root=ROOT, lib=lib, scene=scene, depth=5, w=32, h=24, spp=2, out=str(tmp_path / "py"),
                            config=config, outdir=str(tmp_path / "out"))
    py = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
    assert py.returncode == 0, py.stderr[-3000:]
    host_ppm = (tmp_path / "host.ppm").read_bytes()
    assert host_ppm == (tmp_path / "py.ppm").read_bytes()
    # the entry point's frame of the same samples (fixed_count, no sequence)
    assert host_ppm == (tmp_path / "out" / "tiny.ppm").read_bytes()
    assert (tmp_path / "host.pfm").read_bytes() == (tmp_path / "out" / "tiny.pfm").read_bytes()
