"""The highlight survey's chain walk, on the CPU: the leaf programs it
walks, and the reference's own walkedChain (chain.cpp:507-519, built from
its sources into oracle/_ref) pinned by what the mathematics fixes: where it
settles, a leaf's unit normal halves the directions to the start and to the
lamp, the point lies on the leaf, and over a plane it is the classic mirror
image construction (the start joined to the lamp's image across the plane).
The device's walk is compared with this one in test_gpu_chain.py."""
import math

import numpy as np
import pytest

import kodanuki_amd as k
from tests import chain_cases as CC

# opcodes (kodanuki.h, the reference's Opcode order)
OP_CONSTANT, OP_POINT_X, OP_POINT_Z = 0, 2, 4
OP_ADD, OP_SUBTRACT, OP_MULTIPLY, OP_DIVIDE, OP_NEGATE, OP_SQUARE_ROOT, OP_ABSOLUTE = 8, 9, 10, 11, 12, 13, 14
OP_MINIMUM, OP_MAXIMUM = 21, 22

@pytest.fixture(scope="module")
def scene():
    return k.Scene(CC.MIRRORS)


def _value(steps, values, result, p):
    """A leaf program's value at p (the ops leaf programs use)."""
    r = []
    for s, v in zip(steps, values):
        op, a, b = int(s["op"]), int(s["a"]), int(s["b"])
        if op == OP_CONSTANT:
            x = v
        elif OP_POINT_X <= op <= OP_POINT_Z:
            x = p[op - OP_POINT_X]
        else:
            x = {OP_ADD: lambda: r[a] + r[b], OP_SUBTRACT: lambda: r[a] - r[b], OP_MULTIPLY: lambda: r[a] * r[b],
                 OP_DIVIDE: lambda: r[a] / r[b], OP_NEGATE: lambda: -r[a], OP_SQUARE_ROOT: lambda: math.sqrt(r[a]),
                 OP_ABSOLUTE: lambda: abs(r[a]), OP_MINIMUM: lambda: min(r[a], r[b]),
                 OP_MAXIMUM: lambda: max(r[a], r[b])}[op]()
        r.append(x)
    return r[result]


def test_leaf_programs_are_the_leaves_alone(scene):
    rng = np.random.default_rng(1)
    steps, values, result = scene.leaf_program(0)  # the sphere
    assert all(s["leaf"] == 0 for s in steps)
    for p in rng.uniform(-2, 2, (50, 3)):
        want = np.linalg.norm(p - np.array([-0.8, 0, 0.5])) - 0.5
        assert abs(_value(steps, values, result, p) - want) < 1e-12
    steps, values, result = scene.leaf_program(2)  # the torus, turned about x
    c, s = math.cos(0.4), math.sin(0.4)
    for p in rng.uniform(-2, 2, (50, 3)):
        q = p - np.array([0, 0.9, 0.3])
        local = np.array([q[0], c * q[1] + s * q[2], -s * q[1] + c * q[2]])
        want = math.hypot(math.hypot(local[0], local[1]) - 0.35, local[2]) - 0.1
        assert abs(_value(steps, values, result, p) - want) < 1e-12
    with pytest.raises(k.KodanukiError):
        scene.leaf_program(99)


@pytest.fixture(scope="module")
def reference():
    lib = CC.ref_lib()
    if lib is None:
        pytest.skip("oracle/_ref is not built (make -C oracle ref, with the reference tree)")
    return lib


def _unit(v):
    return v / np.linalg.norm(v)


def test_reference_walk_obeys_the_mirror_law(scene, reference):
    """Settled chains of the sphere and the torus toward a point and a
    sphere lamp: on the sphere, its normal the half vector with start and
    lamp on one side of it (the walk settles on either side of a leaf);
    the axis toward the point; a positive spread."""
    qs = CC.queries()
    qs = qs[(qs["light"] == 0) | (qs["light"] == 2)]
    qs = qs[qs["leaf"] != 1]
    got = CC.ref_chains(reference, scene, qs)
    lights = scene.lights()
    centre = np.array([-0.8, 0, 0.5])
    for q, r in zip(qs, got):
        if not r["found"]:
            continue
        x = r["point"]
        lamp = np.array(lights[int(q["light"])].at, np.float64)
        to_start, to_lamp = _unit(q["from"] - x), _unit(lamp - x)
        half = _unit(to_start + to_lamp)
        if q["leaf"] == 0:
            # the half vector is the normal on whichever side both lie
            n = _unit(x - centre)
            assert abs(np.linalg.norm(x - centre) - 0.5) < 1e-9
            assert min(np.linalg.norm(n - half), np.linalg.norm(n + half)) < 1e-6
            assert np.dot(n, to_start) * np.dot(n, to_lamp) > 0
        assert np.allclose(r["axis"], _unit(x - q["from"]), atol=1e-12)
        assert r["spread"][0] > 0 and r["spread"][2] > 0
    assert got["found"].mean() > 0.3


def test_reference_walk_over_a_plane_is_the_image_construction(scene, reference):
    """Over the box's flat top (z = 0.8, inside its rounded edges) the
    settled point is where the start's line to the lamp's mirror image
    crosses the plane. The starts sit along the mirrored direction of the
    lamp seen from the top's middle, moved a little, so that the point
    stays on the flat part."""
    rng = np.random.default_rng(4)
    lamp = np.array(scene.lights()[0].at, np.float64)
    middle = np.array([0.8, 0.2, 0.8])
    back = (lamp - middle) * np.array([-1, -1, 1])
    qs = []
    for _ in range(40):
        q = np.zeros((), k.CHAIN_QUERY_DTYPE)
        q["from"] = middle + back * rng.uniform(0.6, 1.0) + np.array([rng.uniform(-0.05, 0.05), rng.uniform(-0.05, 0.05), 0])
        q["sideways"] = [[1, 0, 0], [0, 1, 0]]
        q["seed"] = [0.8, 0.2, 0.8]
        q["width"] = 0.05
        q["leaf"] = 1
        q["light"] = 0
        qs.append(q)
    qs = np.array(qs, k.CHAIN_QUERY_DTYPE)
    got = CC.ref_chains(reference, scene, qs)
    assert got["found"].all()
    for q, r in zip(qs, got):
        image = lamp * np.array([1, 1, -1]) + np.array([0, 0, 2 * 0.8])
        s = q["from"]
        t = (0.8 - s[2]) / (image[2] - s[2])
        want = s + (image - s) * t
        assert np.linalg.norm(r["point"] - want) < 1e-9
