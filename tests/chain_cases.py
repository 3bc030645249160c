"""Shared cases of the highlight survey's chain walk: a scene of shiny
leaves (a sphere, a turned rounded box, a turned torus) under one lamp of
each kind, the queries that walk them from points of the floor, and the
reference's own walkedChain (chain.cpp:507-519, built into oracle/_ref) on
the same leaf programs and lamps."""
import ctypes as C
import math
import os

import numpy as np

import kodanuki_amd as k

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF = os.path.join(ROOT, "oracle", "_ref", "libkdn_ref.so")

MIRRORS = """import "standard.sdf";
render { width = 64; height = 48; seed = 1; samples = 4; surfaces = 4; far = 8; }
camera { eye = [0, -4, 1.5]; target = [0, 0, 0.5]; up = [0, 0, 1]; fov = 0.9; }
light point { at = [1.2, -1.0, 2.0]; power = [20, 20, 20]; }
light quad { at = [-1.5, -0.5, 2.2]; along = [0.4, 0, 0]; across = [0, 0.4, 0]; power = [30, 30, 30]; }
light sphere { at = [0.5, 1.5, 1.8]; radius = 0.15; power = [20, 20, 20]; }
light sun { direction = [0.3, -0.4, 1]; angle = 0.05; power = [3, 3, 3]; }
light directional { direction = [-0.5, -0.2, 1]; power = [1, 1, 1]; }
let steel = metal([0.9, 0.9, 0.9], roughness = 0.05);
paint(steel) at([-0.8, 0, 0.5]) sphere(0.5);
paint(steel) at([0.8, 0.2, 0.4]) turn(z, 0.6) box([0.6, 0.4, 0.8], 0.05);
paint(steel) at([0, 0.9, 0.3]) turn(x, 0.4) torus(0.35, 0.1);
paint(matte([0.5, 0.5, 0.5])) at([0, 0, -0.05]) box([6, 6, 0.1]);
"""
# each shiny leaf: its centre and a reach from it to its surface
SHINY = [(0, (-0.8, 0.0, 0.5), 0.5), (1, (0.8, 0.2, 0.4), 0.2), (2, (0.0, 0.9, 0.3), 0.3)]


def queries(n_starts=24, seed=3):
    """Every shiny leaf with every lamp from n_starts floor points: each
    start's sideways directions span the floor; the seed is the leaf's
    centre moved its reach toward the start."""
    rng = np.random.default_rng(seed)
    starts = np.stack([rng.uniform(-2.2, 2.2, n_starts), rng.uniform(-2.2, 2.2, n_starts), np.full(n_starts, 0.001)], 1)
    out = []
    for leaf, centre, reach in SHINY:
        c = np.array(centre)
        for light in range(5):
            for s in starts:
                d = s - c
                q = np.zeros((), k.CHAIN_QUERY_DTYPE)
                q["from"] = s
                q["sideways"] = [[1, 0, 0], [0, 1, 0]]
                q["seed"] = c + d / np.linalg.norm(d) * reach
                q["width"] = 0.05
                q["leaf"] = leaf
                q["light"] = light
                out.append(q)
    return np.array(out, k.CHAIN_QUERY_DTYPE)


def ref_lib():
    if not os.path.exists(REF):
        return None
    lib = C.CDLL(REF)
    P = C.c_void_p
    lib.kr_walked_chain.restype = C.c_int
    lib.kr_walked_chain.argtypes = [P, P, P, C.c_uint32, C.c_uint32, P, P, C.c_double, C.c_int, P, P, P, P, C.c_double, C.c_double, P,
                                    P]
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def lamp_numbers(light):
    """A scene light as the device's lamp table holds it (FP32): its
    fields, the direction made unit, a sun's half-angle cosine."""
    at = np.array(light.at, np.float64)
    along = np.array(light.along, np.float64)
    across = np.array(light.across, np.float64)
    d = np.array(light.direction, np.float64)
    if np.linalg.norm(d) > 0:
        d = (d / np.linalg.norm(d)).astype(np.float32).astype(np.float64)
    cosine = 1.0
    if light.kind in (k.LIGHT_SUN, k.LIGHT_MOON):
        cosine = float(np.float32(math.cos(float(np.float32(light.angle)) / 2)))
    return light.kind, at, d, along, across, float(np.float32(light.radius)), cosine


def ref_chains(lib, scene, qs):
    """The reference's walkedChain of each query, as CHAIN_DTYPE."""
    lights = scene.lights()
    progs = {}
    out = np.zeros(len(qs), k.CHAIN_DTYPE)
    res = np.zeros(21, np.float64)
    for i, q in enumerate(qs):
        leaf = int(q["leaf"])
        if leaf not in progs:
            steps, values, result = scene.leaf_program(leaf)
            ops = np.ascontiguousarray(steps["op"], np.uint8)
            abc = np.ascontiguousarray(np.stack([steps["a"], steps["b"], steps["c"]], 1).astype(np.uint32))
            progs[leaf] = (ops, abc, np.ascontiguousarray(values), len(steps), result)
        ops, abc, values, count, result = progs[leaf]
        kind, at, d, along, across, radius, cosine = lamp_numbers(lights[int(q["light"])])
        args = [np.ascontiguousarray(x, np.float64) for x in (q["from"], q["sideways"].ravel(), at, d, along, across, q["seed"])]
        found = lib.kr_walked_chain(_p(ops), _p(abc), _p(values), count, result, _p(args[0]), _p(args[1]), float(q["width"]), kind,
                                    _p(args[2]), _p(args[3]), _p(args[4]), _p(args[5]), radius, cosine, _p(args[6]), _p(res))
        out[i]["found"] = found
        if found:
            out[i]["point"], out[i]["axis"], out[i]["across"], out[i]["up"] = res[0:3], res[3:6], res[6:9], res[9:12]
            out[i]["spread"] = res[12:15]
            out[i]["motion"] = res[15:21].reshape(2, 3)
    return out
