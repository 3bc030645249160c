"""The C-ABI boundary: the library loads, exports what the header
declares, and reports errors as codes with messages (the reference's
syntax layer returns "line:column: message", source/syntax/evaluator.h:342-369)."""
import os
import re

import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "kodanuki.h")).read()
    return sorted(set(re.findall(r"\b(kdn_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = k.library()
    declared = header_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(k.EXPORTS) == declared


def test_scene_error_is_a_value_with_position():
    with pytest.raises(k.KodanukiError) as e:
        k.Scene("let a = ;")
    assert re.search(r"\d+:\d+: ", str(e.value))
    with pytest.raises(k.KodanukiError):
        k.Scene('import "standard.sdf";\nspindle(1);')


def test_device_calls_fail_loudly_without_a_gpu():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    with pytest.raises(k.KodanukiError):
        k.Device(0)


def test_region_is_exact_power_of_two_box():
    s = k.Scene(scenes.plain(0))
    for depth in (4, 6, 8):
        lower, size = s.region(depth)
        for kk in range(3):
            m, e = np.frexp(size[kk])
            assert m == 0.5
            q = lower[kk] / np.ldexp(size[kk], -depth)
            assert q == np.floor(q)


def test_scene_settings_round_trip():
    s = k.Scene(scenes.plain(3, width=320, height=200, samples=9))
    cam, rs = s.settings()
    assert (rs.width, rs.height, rs.samples) == (320, 200, 9)
    assert s.info.lights == 1
    assert s.info.leaves == 33


def test_struct_layouts_match_the_header(tmp_path):
    """Every struct of include/kodanuki.h has the size its ctypes mirror has
    (a C program built against the header prints them)."""
    import ctypes as C
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("needs gcc")
    names = {"kdn_step": k.Step, "kdn_object": None, "kdn_ray": k.Ray, "kdn_hit": k.Hit, "kdn_material": k.Material,
             "kdn_light": k.Light, "kdn_leaf": k.Leaf, "kdn_node": k.Node, "kdn_camera": k.Camera,
             "kdn_render_settings": k.RenderSettings, "kdn_scene_info": k.SceneInfo, "kdn_octree_stats": k.OctreeStats,
             "kdn_node_record": k.NodeRecord, "kdn_entry_record": k.EntryRecord, "kdn_pass_desc": k.PassDesc,
             "kdn_pass_stats": k.PassStats, "kdn_job_desc": k.JobDesc, "kdn_job_status": k.JobStatus,
             "kdn_serve_stats": k.ServeStats, "kdn_guide_info": k.GuideInfo, "kdn_denoise_frame": k.DenoiseFrame, "kdn_denoise_stats": k.DenoiseStats, "kdn_guide_config": k.GuideConfig, "kdn_guide_query": k.GuideQuery,
             "kdn_guide_sample": k.GuideSample, "kdn_guide_lobe": k.GuideLobe, "kdn_hguide_stats": k.HighlightStats}
    src = tmp_path / "sizes.c"
    src.write_text('#include <stdio.h>\n#include "kodanuki.h"\nint main(void) {\n' +
                   "".join(f'    printf("{n} %zu\\n", sizeof({n}));\n' for n in names) + "    return 0;\n}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    sizes = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for n, t in names.items():
        expect = k.OBJECT_DTYPE.itemsize if t is None else C.sizeof(t)
        assert int(sizes[n]) == expect, (n, sizes[n], expect)
