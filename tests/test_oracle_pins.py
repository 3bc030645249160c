"""Pins of the FP64 oracle's sphere trace and duals (CPU, no GPU).

The FP64 sphere trace (oracle/ref64.c, restating
plugins/geometry/sphere_trace/sphere_trace.cpp:51-228) is the reference the
device's hit distances are held to; these checks fix it independently:
- closed forms: a ray meets a sphere at the smaller root of
  |o + t d - c| = r, and a plane z = 0 at t = -o_z / d_z, with the normal
  of the surface there;
- an independent march of the scene tree's own CSG field (not the compiled
  tape), stepping by the field until it falls under 1e-7 within 4000 steps,
  as the reference's walk checks trace their truth
  (target/tests/walk.cpp:91-109);
- rays from inside a solid: the FP32 twin's march through the octree
  (oracle/twin32.c) and the FP64 trace of the whole scene agree on where
  the ray leaves the solid and through which leaf.

The FP64 duals (or_dual64, restating sdf/numbers.cpp:433-590) against
central differences of the point evaluation, as the reference's number
checks do (target/tests/numbers.cpp:51-91).
"""
import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes
from tests import oracle_py as O
from tests import raysets

_HEAD = """import "standard.sdf";
render { width = 8; height = 8; seed = 1; samples = 1; surfaces = 2; far = 3; }
camera { eye = [0, -4, 0]; target = [0, 0, 0]; up = [0, 0, 1]; fov = 0.8; }
let m = matte([0.5, 0.5, 0.5]);
"""


def _trace(text, rays, depth=6):
    s = k.Scene(text)
    A = O.SceneArrays(s)
    lower, size = s.region(depth)
    hits, d64 = O.sphere_trace64(A, lower, size, rays)
    return s, A, lower, size, hits, d64


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def test_sphere_trace64_meets_a_sphere_at_its_root():
    c = np.array([0.1, 0.2, 0.3])
    r = 0.7
    rng = np.random.default_rng(3)
    o = c + 2.0 * _unit(rng.normal(size=(200, 3)))
    aim = c + 0.6 * r * _unit(rng.normal(size=(200, 3))) * rng.random((200, 1))
    d = _unit(aim - o)
    rays = k.make_rays(o.astype(np.float32), d.astype(np.float32))
    _, _, _, _, hits, d64 = _trace(_HEAD + "paint(m) at([0.1, 0.2, 0.3]) sphere(0.7);\n", rays)
    # the rays as stored (FP32 origin and direction)
    o32 = rays["origin"].astype(np.float64)
    d32 = rays["direction"].astype(np.float64)
    oc = o32 - c
    b = np.sum(oc * d32, axis=1)
    cc = np.sum(oc * oc, axis=1) - r * r
    dd = np.sum(d32 * d32, axis=1)
    t = (-b - np.sqrt(b * b - dd * cc)) / dd
    p = o32 + d32 * t[:, None]
    n = (p - c) / r
    cosine = -np.sum(n * d32, axis=1) / np.sqrt(dd)
    assert np.all(hits["kind"] == k.KDN_HIT_SURFACE)
    # the march stops within 1e-6 of the surface: 1e-6 / cosine along the ray
    assert np.all(np.abs(d64 - t) <= 1e-6 / cosine + 1e-9), np.max(np.abs(d64 - t) * cosine)
    assert np.all(d64 <= t + 1e-12)
    assert np.max(np.abs(hits["normal"] - n)) < 2e-5
    assert np.all(hits["leaf"] == 0)


def test_sphere_trace64_meets_a_plane_and_misses_away_from_it():
    rng = np.random.default_rng(4)
    o = np.column_stack([rng.uniform(-1, 1, 300), rng.uniform(-1, 1, 300), rng.uniform(0.2, 2, 300)])
    d = _unit(np.column_stack([rng.uniform(-1, 1, 300), rng.uniform(-1, 1, 300), rng.uniform(-1, -0.3, 300)]))
    d[::3, 2] *= -1  # a third look away from the plane
    d = _unit(d)
    rays = k.make_rays(o.astype(np.float32), d.astype(np.float32))
    _, _, _, _, hits, d64 = _trace(_HEAD + "paint(m) plane([0, 0, 1], 0);\n", rays)
    o32 = rays["origin"].astype(np.float64)
    d32 = rays["direction"].astype(np.float64)
    down = d32[:, 2] < 0
    t = -o32[:, 2] / d32[:, 2]
    p = o32 + d32 * t[:, None]
    inside = down & np.all(np.abs(p) <= 3, axis=1)  # within the region the trace spans
    assert np.all(hits["kind"][~down] == k.KDN_HIT_MISSED)
    assert np.all(hits["kind"][inside] == k.KDN_HIT_SURFACE)
    cosine = -d32[inside, 2] / np.linalg.norm(d32[inside], axis=1)
    assert np.all(np.abs(d64[inside] - t[inside]) <= 1e-6 / cosine + 1e-9)
    assert np.max(np.abs(hits["normal"][inside] - [0, 0, 1])) < 1e-6


def _independent_march(A, o, d):
    """walk.cpp's truth: step by the scene tree's field from t = 0 until it
    falls under 1e-7 (4000 steps, t < 60)."""
    t = 0.0
    for _ in range(4000):
        if t >= 60:
            break
        v = A.csg_distance(o + d * t)
        if v < 1e-7:
            return t
        t += v
    return None


@pytest.mark.parametrize("name", ["tiny", "plain"])
def test_sphere_trace64_matches_an_independent_march(name):
    text = scenes.tiny() if name == "tiny" else scenes.plain(0)
    s = k.Scene(text)
    A = O.SceneArrays(s)
    lower, size = s.region(6)
    rays = np.concatenate([raysets.camera_rays(s, 60, 11), raysets.random_rays(A, lower, size, 60, 12)])
    hits, d64 = O.sphere_trace64(A, lower, size, rays)
    agree = 0
    checked = 0
    for i, ray in enumerate(rays):
        o = ray["origin"].astype(np.float64)
        d = ray["direction"].astype(np.float64)
        truth = _independent_march(A, o, d)
        # the independent march does not clip to the region: count it a
        # miss where its hit lies outside the region the trace spans
        if truth is not None:
            p = o + d * truth
            if np.any(p < lower - 1e-9) or np.any(p > lower + size + 1e-9):
                truth = None
        found = hits["kind"][i] == k.KDN_HIT_SURFACE
        agree += found == (truth is not None)
        if found and truth is not None:
            n = hits["normal"][i].astype(np.float64)
            cosine = abs(float(np.dot(n, d)))
            if cosine >= 0.05:
                checked += 1
                # both stop short of the surface: the trace within 1e-6,
                # the truth within 1e-7, each along the ray over the cosine
                assert abs(d64[i] - truth) <= 1.2e-6 / cosine, (i, d64[i], truth, cosine)
    assert agree >= len(rays) - 1
    assert checked > len(rays) // 3


def test_inside_rays_twin_and_fp64_agree():
    """Rays from inside a solid leave it where the FP64 trace of the whole
    scene says, through the same leaf (away from seams and ties). The rays
    are within no material, so every surface they come near counts as a
    crossing into another (sphereCrossing, sphere_trace.cpp:176-187)."""
    s = k.Scene(scenes.tiny())
    A = O.SceneArrays(s)
    depth = 6
    lower, size = s.region(depth)
    rays = raysets.random_rays(A, lower, size, 400, 21, inside_share=1.0)
    rays = rays[rays["inside"] == 1]
    assert len(rays) > 200
    tree, _, _, _ = O.build_octree32(A, lower, size, depth)
    try:
        twin = O.trace32(tree, lower, size, depth, A.leaf_material, rays)
    finally:
        O.lib().or_free_octree(tree)
    ref, d64 = O.sphere_trace64(A, lower, size, rays)
    # the same rays leave their solid inside the region (the rest leave it
    # through the region's faces: the ground's solid below the plane)
    assert np.mean(twin["kind"] == ref["kind"]) > 0.99
    both = (twin["kind"] == k.KDN_HIT_SURFACE) & (ref["kind"] == k.KDN_HIT_SURFACE)
    assert both.sum() > 200
    # each stops within its own tolerance of the surface it leaves through
    # (the twin's FP32 march 1e-5 + 2e-6 t, the FP64 trace 1e-6), measured
    # along the ray: over the cosine at which the ray meets that surface
    t = d64[both]
    d = rays["direction"][both].astype(np.float64)
    n = ref["normal"][both].astype(np.float64)
    cosine = np.maximum(np.abs(np.sum(d * n, axis=1)), 1e-3)
    gap = np.abs(twin["distance"][both].astype(np.float64) - t)
    bound = (1e-5 + 2e-6 * t + 1e-6) / cosine + 1e-6
    assert np.mean(gap <= bound) > 0.99, np.max(gap / bound)
    assert np.mean(twin["leaf"][both] == ref["leaf"][both]) > 0.98


def _every_opcode_program():
    S = k.STEP_DTYPE
    prog = [
        (2, 0, 0, 0, 0.0), (3, 0, 0, 0, 0.0), (4, 0, 0, 0, 0.0),
        (0, 0, 0, 0, 0.7), (0, 0, 0, 0, 2.0), (0, 0, 0, 0, 0.3),
        (8, 0, 1, 0, 0.0), (9, 6, 2, 0, 0.0), (10, 7, 3, 0, 0.0),      # 6 add 7 sub 8 mul
        (11, 8, 4, 0, 0.0), (12, 9, 0, 0, 0.0), (14, 10, 0, 0, 0.0),   # 9 div 10 neg 11 abs
        (8, 11, 4, 0, 0.0), (13, 12, 0, 0, 0.0), (16, 7, 0, 0, 0.0),   # 12 add 13 sqrt 14 sin
        (17, 8, 0, 0, 0.0), (19, 1, 0, 0, 0.0), (20, 12, 4, 0, 0.0),   # 15 cos 16 atan2 17 pow
        (21, 13, 14, 0, 0.0), (22, 15, 10, 0, 0.0), (23, 18, 19, 5, 0.0),  # 18 min 19 max 20 smin
        (24, 7, 0, 0, 0.0), (8, 20, 21, 0, 0.0), (8, 22, 16, 0, 0.0),  # 21 pos2 22 23 adds
        (8, 23, 17, 0, 0.0),
    ]
    steps = np.zeros(len(prog), S)
    values = np.zeros(len(prog), np.float64)
    for i, (op, a, b, c, v) in enumerate(prog):
        steps[i] = (op, 0, a, b, c, np.float32(v), 0)
        values[i] = v
    return steps, values, len(prog) - 1


def _dual64(steps, values, result, p):
    regs = np.zeros(4 * len(steps), np.float64)
    g = np.zeros(3, np.float64)
    p = np.ascontiguousarray(p, np.float64)
    v = O.lib().or_dual64(O.ptr(steps), O.ptr(values), len(steps), result, O.ptr(p), O.ptr(g), O.ptr(regs))
    return v, g


def _check_gradients(steps, values, result, points):
    h = 1e-5
    worst = 0.0
    for p in points:
        v, g = _dual64(steps, values, result, p)
        assert abs(v - O.eval64(steps, values, result, p)) <= 1e-12 * (1 + abs(v))
        for i in range(3):
            up, down = p.copy(), p.copy()
            up[i] += h
            down[i] -= h
            diff = (O.eval64(steps, values, result, up) - O.eval64(steps, values, result, down)) / (2 * h)
            worst = max(worst, abs(g[i] - diff) / (1 + abs(diff)))
    return worst


def test_dual64_every_opcode_against_central_differences():
    steps, values, result = _every_opcode_program()
    rng = np.random.default_rng(5)
    # away from the kinks of abs, min, max and the band edges, where no
    # difference quotient is a derivative
    points = []
    while len(points) < 200:
        p = rng.uniform(0.2, 1.5, 3)
        regs = np.zeros(len(steps), np.float64)
        O.lib().or_eval64(O.ptr(steps), O.ptr(values), len(steps), result, O.ptr(p), O.ptr(regs))
        if min(abs(regs[13] - regs[14]), abs(regs[15] - regs[10]), abs(abs(regs[18] - regs[19]) - 0.3),
               abs(regs[7])) > 1e-3:
            points.append(p)
    assert _check_gradients(steps, values, result, points) < 1e-6


@pytest.mark.parametrize("name", ["tiny", "plain", "production"])
def test_dual64_scene_programs_against_central_differences(name):
    text = {"tiny": scenes.tiny(), "plain": scenes.plain(0), "production": scenes.production(0, count=16)}[name]
    s = k.Scene(text)
    A = O.SceneArrays(s)
    lower, size = s.region(6)
    rng = np.random.default_rng(6)
    worst = 0.0
    for j in range(len(A.objects)):
        st, va, res = A.program(j)
        points = [lower + rng.random(3) * size for _ in range(40)]
        worst = max(worst, _check_gradients(st, va, res, points))
    # the scenes' fields are smooth away from measure-zero seams; a point
    # within 1e-5 of one is a rare event at these counts
    assert worst < 1e-4
