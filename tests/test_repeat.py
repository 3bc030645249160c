"""The reference's repeat (syntax/evaluator.cpp:2457-2577) in the loader, and
the stress scene built on it.

repeat(count = [nx, ny, nz], step = [sx, sy, sz]) { ... } emits the block's
shapes, then copies at every other place of the grid (z outermost, x
innermost), offset along the frame's axes; at the top level each copy is an
object of its own, inside a join they are merged. Checked against the same
shapes written out one by one, and the reference's refusals.
"""
import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes
from tests import oracle_py as O

HEAD = 'import "standard.sdf";\nrender { width = 32; height = 32; }\ncamera { eye = [0, -5, 2]; target = [0, 0, 0]; up = [0, 0, 1]; fov = 0.8; }\n'


def _fields_equal(a, b, lower, size, n=400):
    A, B = O.SceneArrays(a), O.SceneArrays(b)
    rng = np.random.default_rng(3)
    for _ in range(n):
        p = lower + rng.random(3) * size
        fa, fb = A.csg_distance(p), B.csg_distance(p)
        assert abs(fa - fb) <= 1e-12 * (1 + abs(fb)), (p, fa, fb)
        assert abs(A.tape_distance64(p) - fb) <= 1e-12 * (1 + abs(fb))


def test_repeat_is_the_copies_written_out():
    rep = k.Scene(HEAD + "let a = matte([0.5, 0.5, 0.5]);\nat([0.2, 0, 0.1]) turn(z, 0.7) repeat(count = [3, 2, 2], step = [0.8, 0.5, 0.3]) "
                         "{ paint(a) sphere(0.15); paint(a) at([0.1, 0, 0]) box([0.1, 0.2, 0.1], 0.01); }\n")
    c, s = np.cos(0.7), np.sin(0.7)
    lines = []
    for kk in range(2):
        for j in range(2):
            for i in range(3):
                lx, ly, lz = 0.8 * i, 0.5 * j, 0.3 * kk
                x, y, z = 0.2 + c * lx - s * ly, s * lx + c * ly, 0.1 + lz
                lines.append(f"at([{float(x)!r}, {float(y)!r}, {float(z)!r}]) turn(z, 0.7) {{ paint(a) sphere(0.15); paint(a) at([0.1, 0, 0]) box([0.1, 0.2, 0.1], 0.01); }}")
    out = k.Scene(HEAD + "let a = matte([0.5, 0.5, 0.5]);\n" + "\n".join("with " + ln for ln in lines) + "\n")
    assert rep.info.leaves == out.info.leaves == 24
    assert rep.info.objects == out.info.objects == 24
    lower, size = rep.region(6)
    _fields_equal(rep, out, lower, size)


def test_repeat_inside_a_join_is_merged():
    s = k.Scene(HEAD + "merge(0.05) { repeat(count = [4, 1, 1], step = [0.3, 0, 0]) { sphere(0.1); } box([1, 1, 0.1], 0); }\n")
    assert s.info.leaves == 5 and s.info.objects == 1


@pytest.mark.parametrize("text,message", [
    ("repeat(count = [2, 1, 1]) { sphere(0.1); }", "repeat takes count and step"),
    ("repeat(count = [2.5, 1, 1], step = [1, 0, 0]) { sphere(0.1); }", "whole number of one or more"),
    ("repeat(count = [0, 1, 1], step = [1, 0, 0]) { sphere(0.1); }", "whole number of one or more"),
    ("repeat(count = [2, 1, 1], step = [1, 0, 0]) { }", "emits no shape"),
])
def test_repeat_refusals(text, message):
    with pytest.raises(k.KodanukiError, match=message):
        k.Scene(HEAD + text + "\n")


def test_stress_scene_shapes():
    """The full stress scene: 4096 primitives and a 64 x 64 canopy of
    bubbles (a thin shell and a nested core each); glass and its matte
    variant differ in materials only."""
    g, m = k.Scene(scenes.stress(0)), k.Scene(scenes.stress(0, glass=False))
    assert g.info.leaves == m.info.leaves == 4096 + 64 * 64 * 3 + 1
    assert g.info.steps == m.info.steps
    kinds = [mm.kind for mm in g.materials()]
    assert kinds.count(k.GLASS) >= 2
    assert k.GLASS not in [mm.kind for mm in m.materials()]
    lower, size = g.region(10)
    assert lower[2] <= 0 and lower[2] + size[2] >= 3.2 + 0.08
