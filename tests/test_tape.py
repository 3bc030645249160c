"""The tape and its interval cut on the CPU (oracle against closed forms
and invariants; no GPU).

- A1: the compiled object programs give the field of the scene tree,
  evaluated straight from its CSG nodes (source/sdf/scene.cpp:36-75).
- A4: an interval contains the point values of its box (as
  target/tests/numbers.cpp:105-114); the FP32 outward intervals contain the
  reference's FP64 intervals.
- A5-A7: a cut program equals the whole program inside its cell, Empty and
  Filled cells agree with point samples (target/tests/tree.cpp:165-223), and
  the twin octree is deterministic.
"""
import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes
from tests import oracle_py as O


@pytest.fixture(scope="module", params=["tiny", "plain", "production", "stress"])
def scene(request):
    text = {"tiny": scenes.tiny(), "plain": scenes.plain(0), "production": scenes.production(0, count=64),
            "stress": scenes.stress(0, count=64, grid=4)}[request.param]
    s = k.Scene(text)
    return s, O.SceneArrays(s)


def test_tape_equals_scene_tree(scene):
    s, A = scene
    lower, size = s.region(6)
    rng = np.random.default_rng(7)
    for _ in range(300):
        p = lower + rng.random(3) * size
        a, b = A.tape_distance64(p), A.csg_distance(p)
        assert abs(a - b) <= 1e-12 * (1 + abs(b)), (p, a, b)


def test_every_opcode_interval_contains_points():
    """A program that uses every opcode family (numbers.cpp test shape)."""
    S = k.STEP_DTYPE
    prog = [
        (2, 0, 0, 0, 0.0), (3, 0, 0, 0, 0.0), (4, 0, 0, 0, 0.0),          # 0-2 x y z
        (0, 0, 0, 0, 0.7), (0, 0, 0, 0, 2.0), (0, 0, 0, 0, 0.05),         # 3-5 consts
        (8, 0, 1, 0, 0.0), (9, 6, 2, 0, 0.0), (10, 7, 3, 0, 0.0),         # 6 add 7 sub 8 mul
        (11, 8, 4, 0, 0.0), (12, 9, 0, 0, 0.0), (14, 10, 0, 0, 0.0),      # 9 div 10 neg 11 abs
        (13, 11, 0, 0, 0.0), (15, 6, 0, 0, 0.0), (16, 7, 0, 0, 0.0),      # 12 sqrt 13 floor 14 sin
        (17, 8, 0, 0, 0.0), (18, 9, 0, 0, 0.0), (19, 1, 0, 0, 0.0),       # 15 cos 16 sign 17 atan2(y,x)
        (20, 11, 4, 0, 0.0), (21, 12, 14, 0, 0.0), (22, 15, 10, 0, 0.0),  # 18 pow 19 min 20 max
        (23, 19, 20, 5, 0.0), (24, 7, 0, 0, 0.0), (25, 0, 1, 0, 0.0),     # 21 smin 22 pos2 23 less
        (26, 1, 2, 0, 0.0), (27, 16, 16, 0, 0.0), (28, 23, 24, 0, 0.0),   # 24 le 25 eq 26 ne
        (29, 23, 26, 0, 0.0), (30, 24, 25, 0, 0.0), (31, 27, 0, 0, 0.0),  # 27 and 28 or 29 not
        (32, 28, 21, 22, 0.0), (8, 30, 18, 0, 0.0), (8, 31, 17, 0, 0.0),  # 30 select 31 32 adds
        (8, 32, 13, 0, 0.0),
    ]
    steps = np.zeros(len(prog), S)
    values = np.zeros(len(prog), np.float64)
    for i, (op, a, b, c, v) in enumerate(prog):
        f = np.float32(v)
        flags = 0 if float(f) == v else (1 if float(f) < v else 2)
        steps[i] = (op, flags, a, b, c, f, 0)
        values[i] = v
    rng = np.random.default_rng(11)
    for _ in range(40):
        c = rng.uniform(-2, 2, 3)
        w = rng.uniform(0.01, 1.0, 3)
        box = np.array([c[0], c[0] + w[0], c[1], c[1] + w[1], c[2], c[2] + w[2]]).astype(np.float32).astype(np.float64)
        lo64, hi64 = O.interval64(steps, values, box)
        lo32, hi32 = O.interval32(steps, box)
        # FP32 outward intervals contain the reference's FP64 intervals
        ok = (lo32.astype(np.float64) <= lo64 + 1e-12 * np.abs(lo64)) | np.isnan(lo64)
        ok &= (hi32.astype(np.float64) >= hi64 - 1e-12 * np.abs(hi64)) | np.isnan(hi64)
        assert ok.all(), np.nonzero(~ok)
        # both contain 50 point values of the box (numbers.cpp:105-114)
        for _ in range(50):
            p = box[0::2] + rng.random(3) * (box[1::2] - box[0::2])
            regs = np.zeros(len(steps))
            O.lib().or_eval64(O.ptr(steps), O.ptr(values), len(steps), 0, O.ptr(p), O.ptr(regs))
            fin = np.isfinite(regs)
            assert np.all((lo64[fin] - 1e-12 <= regs[fin]) & (regs[fin] <= hi64[fin] + 1e-12))
            assert np.all((lo32[fin] <= regs[fin]) & (regs[fin] <= hi32[fin]))


def test_fp32_intervals_contain_fp64_on_scene(scene):
    s, A = scene
    lower, size = s.region(6)
    rng = np.random.default_rng(3)
    for j in range(min(len(A.objects), 12)):
        st, va, res = A.program(j)
        for _ in range(20):
            c = lower + rng.random(3) * size
            w = size * rng.uniform(0.001, 0.2)
            box = np.array([c[0], c[0] + w[0], c[1], c[1] + w[1], c[2], c[2] + w[2]], np.float64)
            box32 = box.astype(np.float32).astype(np.float64)
            lo64, hi64 = O.interval64(st, va, box32)
            lo32, hi32 = O.interval32(st, box32.astype(np.float32))
            assert np.all(lo32 <= lo64 + 1e-12 * np.abs(lo64))
            assert np.all(hi32 >= hi64 - 1e-12 * np.abs(hi64))


def _leaf_programs(nodes, entries, pool, node):
    out = []
    for e in range(node["entry_first"], node["entry_first"] + node["entry_count"]):
        at = entries[e]["program"]
        h = pool[at]
        n = int(h["a"]) | (int(h["b"]) << 16)
        out.append((entries[e]["object"], np.ascontiguousarray(pool[at + 1: at + 1 + n]), int(h["c"])))
    return out


def _box(lower, size, node):
    cell = size * np.ldexp(1.0, -int(node["level"]))
    lo = lower + np.array([node["x"], node["y"], node["z"]]) * cell
    return lo, lo + cell


def test_cut_equals_whole_inside_cell_and_occupancy_agrees(scene):
    s, A = scene
    depth = 5
    lower, size = s.region(depth)
    tree, nodes, entries, pool = O.build_octree32(A, lower, size, depth)
    try:
        rng = np.random.default_rng(5)
        picks = rng.choice(len(nodes), size=min(400, len(nodes)), replace=False)
        for i in picks:
            n = nodes[i]
            lo, hi = _box(lower.astype(np.float64), size.astype(np.float64), n)
            progs = _leaf_programs(nodes, entries, pool, n)
            for _ in range(8):
                p = lo + rng.random(3) * (hi - lo)
                field = A.tape_distance64(p)
                if n["kind"] == k.KIND_EMPTY:
                    assert field > -1e-9, (i, field)
                elif n["kind"] == k.KIND_FILLED:
                    assert field < 1e-9, (i, field)
                for obj, st, res in progs:
                    # the cut program, evaluated in FP64, equals its object's
                    # whole program with the same FP32 constants
                    whole_st, _, whole_res = A.program(obj)
                    whole = O.eval64(whole_st, whole_st["value"].astype(np.float64), whole_res, p)
                    cut = O.eval64(st, st["value"].astype(np.float64), res, p)
                    assert abs(cut - whole) <= 1e-9 * (1 + abs(whole)), (i, obj, cut, whole)
                if n["kind"] == k.KIND_SURFACE:
                    # dropping objects only raises the field, and where the
                    # field is not positive the kept objects decide it
                    # (cut.cpp:446-461): the surface inside the cell is theirs
                    kept = min(O.eval64(st, st["value"].astype(np.float64), res, p) for _, st, res in progs)
                    whole = min(O.eval64(*(lambda t: (t[0], t[0]["value"].astype(np.float64), t[2]))(A.program(j)), p)
                                for j in range(len(A.objects)))
                    assert kept >= whole - 1e-9 * (1 + abs(whole)), (i, kept, whole)
                    if whole <= 0:
                        assert abs(kept - whole) <= 1e-9 * (1 + abs(whole)), (i, kept, whole)
    finally:
        O.lib().or_free_octree(tree)


def test_twin_octree_is_deterministic_and_cut_shrinks():
    s = k.Scene(scenes.plain(1))
    A = O.SceneArrays(s)
    lower, size = s.region(5)
    t1, n1, e1, p1 = O.build_octree32(A, lower, size, 5)
    t2, n2, e2, p2 = O.build_octree32(A, lower, size, 5)
    try:
        assert n1.tobytes() == n2.tobytes() and e1.tobytes() == e2.tobytes() and p1.tobytes() == p2.tobytes()
        # small cells cut their programs far below the whole (tree.cpp:190-200)
        whole = int(A.objects["count"].max())
        leaf = n1[n1["kind"] == k.KIND_SURFACE]
        lens = [int(p1[e1[e]["program"]]["a"]) for n in leaf for e in range(n["entry_first"], n["entry_first"] + n["entry_count"])]
        assert np.median(lens) < whole / 4
    finally:
        O.lib().or_free_octree(t1)
        O.lib().or_free_octree(t2)
