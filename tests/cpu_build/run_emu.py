"""Drive the bounds-checked CPU build of the library through the C ABI.

Runs the GPU test session's sequence of calls (octree builds of the tiny,
plain and production scenes, extend batches, the furnace renders, the
progressive and oracle renders, a counted and timed production pass, a
mix of jobs through the serving loop) on
`libkodanuki_emu.so`, the library's own sources built by g++ with
AddressSanitizer against a CPU stand-in of the HIP runtime. Any out-of-bounds
access in a kernel or a launch aborts with a report. Start it with the
sanitizer runtime preloaded (tests/test_cpu_emulation.py does).
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import kodanuki_amd as k  # noqa: E402
from kodanuki_amd import scenes  # noqa: E402
from tests import oracle_py as O  # noqa: E402
from tests import raysets  # noqa: E402


def say(msg):
    print(msg, flush=True)


def furnace(material):
    return f"""import "standard.sdf";
render {{ width = 32; height = 32; seed = 3; samples = 64; surfaces = 64; far = 2; }}
camera {{ eye = [0, -3, 0]; target = [0, 0, 0]; up = [0, 0, 1]; fov = 0.6; }}
sky gradient {{ zenith = [1, 1, 1]; horizon = [1, 1, 1]; ground = [1, 1, 1]; }}
let m = {material};
paint(m) sphere(0.5);
paint(m) at([0.3, 0.3, 0.3]) box([0.4, 0.4, 0.4], 0.05);
"""


def geometry(dev, quick):
    cases = [("tiny", scenes.tiny(), 5, 2048), ("plain", scenes.plain(0), 6, 4096)]
    cases.append(("production", scenes.production(0), 6 if quick else 8, 1024))
    for name, text, depth, count in cases:
        s = k.Scene(text)
        A = O.SceneArrays(s)
        lower, size = s.region(depth)
        dev.upload(s)
        st = dev.build_octree(lower, size, depth)
        nodes, entries, pool = dev.octree()
        tree, tn, te, tp = O.build_octree32(A, lower, size, depth)
        same = nodes.tobytes() == tn.tobytes() and entries.tobytes() == te.tobytes() and pool.tobytes() == tp.tobytes()
        say(f"[octree] {name} d{depth}: {len(nodes)} nodes {len(entries)} entries {len(pool)} steps, twin equal {same}")
        assert same
        cam = raysets.camera_rays(s, count // 2, 42)
        rnd = raysets.random_rays(A, lower, size, count - count // 2, 43, inside_share=0.2)
        rays = np.concatenate([cam, rnd])
        hits = dev.extend(rays)
        twin = O.trace32(tree, lower, size, depth, A.leaf_material, rays)
        O.lib().or_free_octree(tree)
        same = hits.tobytes() == twin.tobytes()
        say(f"[extend] {name}: {len(rays)} rays, twin equal {same}")
        assert same
        outside = rays[rays["inside"] == 0][:256]
        hits = dev.extend(outside)
        ref, d64 = O.sphere_trace64(A, lower, size, outside)
        both = (hits["kind"] == k.KDN_HIT_SURFACE) & (ref["kind"] == k.KDN_HIT_SURFACE)
        say(f"[fp64] {name}: {len(outside)} rays, both hit {both.sum()}")


def render(dev, text, depth, w, h, spp, surfaces=None, first=0, keep=False, flags=0):
    s = k.Scene(text)
    cam, rs = s.settings()
    lower, size = s.region(depth)
    dev.upload(s)
    dev.build_octree(lower, size, depth)
    st = dev.render_pass(w, h, spp, first_sample=first, seed=rs.seed, surfaces=surfaces or rs.surfaces, keep=keep, flags=flags)
    rgb, sq, n = dev.accumulation(w, h)
    return rgb / n[..., None], st


def integrator(dev, quick):
    spp = 8 if quick else 64
    for m in ["matte([1, 1, 1])", "metal([1, 1, 1], roughness = 0.6)", "glass(refraction = 1.5, dispersion = 0.008)"]:
        mean, st = render(dev, furnace(m), 5, 32, 32, spp)
        say(f"[furnace] {m}: mean {mean.mean():.4f} segments {st.segments}")
    a, _ = render(dev, scenes.tiny(), 5, 24, 18, 16)
    b, _ = render(dev, scenes.tiny(), 5, 24, 18, 8)
    b, _ = render(dev, scenes.tiny(), 5, 24, 18, 8, first=8, keep=True)
    say(f"[progressive] largest difference {np.abs(a - b).max():.2e}")
    assert np.allclose(a, b, rtol=0, atol=1e-9)
    mean, st = render(dev, scenes.tiny(), 5, 24, 18, spp, surfaces=5)
    say(f"[tiny] mean {mean.mean(axis=(0, 1))} segments {st.segments}")
    s = k.Scene(scenes.tiny())
    lower, size = s.region(5)
    cpu = O.render64(s, O.SceneArrays(s), lower, size, 24, 18, 4, seed=s.settings()[1].seed, surfaces=5)
    say(f"[oracle] mean {(cpu / 4).mean(axis=(0, 1))}")


def production(dev, quick):
    text = scenes.production(0, width=1920, height=1080, samples=64, surfaces=13)
    w, h = (48, 27) if quick else (96, 54)
    for flags in (k.PASS_COUNT, k.PASS_TIME):
        mean, st = render(dev, text, 6 if quick else 8, w, h, 1, flags=flags)
        say(f"[production] flags {flags}: segments {st.segments} shadows {st.shadow_rays} flops {st.trace_flops:.3g}")


def serving(dev):
    """Two jobs alone, then together with a third arriving mid-run, on a
    round of 65536 paths (the largest job takes two slices)."""
    plain, tiny = k.Scene(scenes.plain(0)), k.Scene(scenes.tiny())
    jobs = [(plain, 80, 60, 16, 3, 6, 6), (tiny, 31, 23, 5, 4, 8, 5), (plain, 20, 12, 2, 5, 4, 6)]
    alone = []
    with k.Server(dev, max_paths=1 << 16) as server:
        for scene, w, h, spp, seed, surfaces, depth in jobs:
            j = server.submit(scene, w, h, spp, seed=seed, surfaces=surfaces, depth=depth)
            server.run([j])
            alone.append(server.result(j))
    with k.Server(dev, max_paths=1 << 16) as server:
        ids = [server.submit(s, w, h, spp, seed=seed, surfaces=sf, depth=d) for s, w, h, spp, seed, sf, d in jobs[:2]]
        server.step()
        ids.append(server.submit(*jobs[2][:4], seed=jobs[2][4], surfaces=jobs[2][5], depth=jobs[2][6]))
        server.run()
        for i, j in enumerate(ids):
            same = all(a.tobytes() == b.tobytes() for a, b in zip(server.result(j), alone[i]))
            say(f"[serving] job {i}: equal alone and in the mix {same}")
            assert same


def highlight(dev):
    """The plain guide's survey and pixel guides of the highlight room, and
    a guided render beside a plain one."""
    s = k.Scene(scenes.highlight_room())
    lower, size = s.region(5)
    dev.upload(s)
    dev.build_octree(lower, size, 5)
    st = dev.highlight_guides(40, 30)
    say(f"[highlight] points {st.survey_points} highlights {st.highlights} guided pixels {st.guided_pixels} "
        f"lumps {st.mean_lumps:.2f} share {st.mean_share:.3f}")
    g = dev.highlight_guide_at(30, 10)
    say(f"[highlight] pixel (30, 10): count {g[0]:.0f} material {g[1]:.0f} share {g[2]:.3f}")
    a = dev.render_pass(40, 30, 4, seed=1, surfaces=5, flags=k.PASS_HIGHLIGHT)
    say(f"[highlight] guided pass segments {a.segments}")


def main():
    quick = "--quick" in sys.argv
    O.PATH = os.path.join(HERE, "liboracle_asan.so")
    lib = C.CDLL(os.path.join(HERE, "libkodanuki_emu.so"))
    k._bind(lib)
    dev = k.Device(0)
    geometry(dev, quick)
    integrator(dev, quick)
    production(dev, quick)
    serving(dev)
    highlight(dev)
    dev.close()
    say("[emu] done")


if __name__ == "__main__":
    main()
