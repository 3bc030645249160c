"""ctypes access to oracle/liboracle.so — the CPU checker of the tests.

The oracle is test infrastructure: the tests call it to produce the
expected answers the GPU library is compared against.
"""
import ctypes as C
import os

import numpy as np

import kodanuki_amd as k

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "liboracle.so")
P = C.c_void_p


class Octree(C.Structure):
    _fields_ = [("nodes", P), ("entries", P), ("pool", P), ("n_nodes", C.c_uint64), ("n_entries", C.c_uint64),
                ("n_pool", C.c_uint64), ("cap_nodes", C.c_uint64), ("cap_entries", C.c_uint64), ("cap_pool", C.c_uint64)]


class Scene64(C.Structure):
    _fields_ = [("steps", P), ("values", P), ("objects", P), ("object_count", C.c_uint32), ("leaf_material", P),
                ("lower", C.c_double * 3), ("size", C.c_double * 3)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(PATH):
            raise RuntimeError(f"{PATH} is not built: make -C oracle")
        L = C.CDLL(PATH)
        L.or_eval64.restype = C.c_double
        L.or_eval64.argtypes = [P, P, C.c_uint32, C.c_uint32, P, P]
        L.or_interval64.argtypes = [P, P, C.c_uint32, P, P, P]
        L.or_dual64.restype = C.c_double
        L.or_dual64.argtypes = [P, P, C.c_uint32, C.c_uint32, P, P, P]
        L.or_scene_distance.restype = C.c_double
        L.or_scene_distance.argtypes = [P, P, P, P, C.c_uint32, P]
        L.or_interval32.argtypes = [P, C.c_uint32, P, P, P]
        L.or_eval32.restype = C.c_float
        L.or_eval32.argtypes = [P, C.c_uint32, C.c_uint32, P, P]
        L.or_build_octree32.restype = C.c_int
        L.or_build_octree32.argtypes = [P, P, C.c_uint32, P, P, C.c_int, C.POINTER(Octree)]
        L.or_free_octree.argtypes = [C.POINTER(Octree)]
        L.or_trace32.argtypes = [C.POINTER(Octree), P, P, C.c_int, P, P, C.c_uint32, P]
        L.or_sphere_trace64.argtypes = [C.POINTER(Scene64), P, C.c_uint32, P, P]
        L.or_mixed_bits.restype = C.c_uint64
        L.or_mixed_bits.argtypes = [C.c_uint64]
        L.or_hashed_number.restype = C.c_double
        L.or_hashed_number.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.or_sobol_number.restype = C.c_double
        L.or_sobol_number.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        L.or_escape_of.argtypes = [C.c_double, P, P]
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(P)


class SceneArrays:
    """A scene's tape and tree as numpy arrays, for the oracle calls."""

    def __init__(self, scene):
        self.scene = scene
        self.steps, self.objects = scene.tape()
        self.values = scene.tape_values()
        leaves = scene.leaves()
        self.leaves = (k.Leaf * max(1, len(leaves)))(*leaves)
        nodes, tops, children = scene.nodes()
        self.nodes = (k.Node * max(1, len(nodes)))(*nodes)
        self.tops = np.ascontiguousarray(tops, np.uint32)
        self.children = np.ascontiguousarray(children if len(children) else np.zeros(1, np.uint32), np.uint32)
        mats = scene.materials()
        nmat = len(mats)
        self.leaf_material = np.array([l.material if l.material != k.NO_LEAF else nmat for l in leaves] or [nmat], np.uint32)

    def program(self, j):
        o = self.objects[j]
        return (np.ascontiguousarray(self.steps[o["first"]:o["first"] + o["count"]]),
                np.ascontiguousarray(self.values[o["first"]:o["first"] + o["count"]]), int(o["result"]))

    def csg_distance(self, p):
        p = np.ascontiguousarray(p, np.float64)
        return lib().or_scene_distance(C.cast(self.leaves, P), C.cast(self.nodes, P), ptr(self.children), ptr(self.tops),
                                       len(self.tops), ptr(p))

    def tape_distance64(self, p):
        p = np.ascontiguousarray(p, np.float64)
        best = 1e30
        for j in range(len(self.objects)):
            st, va, res = self.program(j)
            regs = np.zeros(len(st), np.float64)
            best = min(best, lib().or_eval64(ptr(st), ptr(va), len(st), res, ptr(p), ptr(regs)))
        return best


def eval64(steps, values, result, p):
    regs = np.zeros(len(steps), np.float64)
    p = np.ascontiguousarray(p, np.float64)
    return lib().or_eval64(ptr(steps), ptr(values), len(steps), result, ptr(p), ptr(regs))


def interval64(steps, values, box):
    lo = np.zeros(len(steps), np.float64)
    hi = np.zeros(len(steps), np.float64)
    box = np.ascontiguousarray(box, np.float64)
    lib().or_interval64(ptr(steps), ptr(values), len(steps), ptr(box), ptr(lo), ptr(hi))
    return lo, hi


def interval32(steps, box):
    lo = np.zeros(len(steps), np.float32)
    hi = np.zeros(len(steps), np.float32)
    box = np.ascontiguousarray(box, np.float32)
    lib().or_interval32(ptr(steps), len(steps), ptr(box), ptr(lo), ptr(hi))
    return lo, hi


def eval32(steps, result, p):
    regs = np.zeros(len(steps), np.float32)
    p = np.ascontiguousarray(p, np.float32)
    return lib().or_eval32(ptr(steps), len(steps), result, ptr(p), ptr(regs))


def build_octree32(arrays, lower, size, depth):
    """The FP32 twin octree: (nodes, entries, pool) numpy arrays."""
    t = Octree()
    lo = np.ascontiguousarray(lower, np.float32)
    sz = np.ascontiguousarray(size, np.float32)
    rc = lib().or_build_octree32(ptr(arrays.steps), ptr(arrays.objects), len(arrays.objects), ptr(lo), ptr(sz), depth, C.byref(t))
    assert rc == 0
    nodes = np.ctypeslib.as_array(C.cast(t.nodes, C.POINTER(C.c_uint8)), (t.n_nodes * k.NODE_DTYPE.itemsize,)).view(k.NODE_DTYPE).copy()
    entries = np.ctypeslib.as_array(C.cast(t.entries, C.POINTER(C.c_uint8)), (max(1, t.n_entries) * k.ENTRY_DTYPE.itemsize,)).view(k.ENTRY_DTYPE)[: t.n_entries].copy()
    pool = np.ctypeslib.as_array(C.cast(t.pool, C.POINTER(C.c_uint8)), (t.n_pool * k.STEP_DTYPE.itemsize,)).view(k.STEP_DTYPE).copy()
    return t, nodes, entries, pool


def trace32(tree, lower, size, depth, leaf_material, rays):
    hits = np.zeros(len(rays), k.HIT_DTYPE)
    rays = np.ascontiguousarray(rays, k.RAY_DTYPE)
    lo = np.ascontiguousarray(lower, np.float32)
    sz = np.ascontiguousarray(size, np.float32)
    lib().or_trace32(C.byref(tree), ptr(lo), ptr(sz), depth, ptr(leaf_material), ptr(rays), len(rays), ptr(hits))
    return hits


def sphere_trace64(arrays, lower, size, rays):
    s = Scene64()
    s.steps = ptr(arrays.steps).value
    s.values = ptr(arrays.values).value
    s.objects = ptr(arrays.objects).value
    s.object_count = len(arrays.objects)
    s.leaf_material = ptr(arrays.leaf_material).value
    for i in range(3):
        s.lower[i] = float(lower[i])
        s.size[i] = float(size[i])
    rays = np.ascontiguousarray(rays, k.RAY_DTYPE)
    hits = np.zeros(len(rays), k.HIT_DTYPE)
    d64 = np.zeros(len(rays), np.float64)
    lib().or_sphere_trace64(C.byref(s), ptr(rays), len(rays), ptr(hits), ptr(d64))
    return hits, d64


class RenderScene(C.Structure):
    _fields_ = [("geometry", Scene64), ("materials", P), ("material_count", C.c_uint32), ("lights", P), ("light_count", C.c_uint32),
                ("camera", k.Camera), ("sky_zenith", C.c_float * 3), ("sky_horizon", C.c_float * 3), ("sky_ground", C.c_float * 3)]


def render64(scene, arrays, lower, size, width, height, spp, first=0, seed=1, surfaces=7, min_depth=4, window=None, threads=None,
             sobol=False):
    """The FP64 path tracer over a window of the image: rgb sums (h, w, 3);
    with sobol, an eye path's first 28 coordinates off the Sobol member."""
    L = lib()
    L.or_render64_sequenced.argtypes = [C.POINTER(RenderScene), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint32,
                                        C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, P, C.c_int]
    x0, y0, x1, y1 = window or (0, 0, width, height)
    mats = scene.materials()
    grey = k.Material()
    grey.kind = 0
    grey.albedo[:] = [0.5, 0.5, 0.5]
    grey.roughness = 1
    grey.strength = 1
    mats = mats + [grey]
    marr = (k.Material * len(mats))(*mats)
    lights = scene.lights()
    larr = (k.Light * max(1, len(lights)))(*lights)
    cam, rs = scene.settings()
    R = RenderScene()
    g = R.geometry
    g.steps = ptr(arrays.steps).value
    g.values = ptr(arrays.values).value
    g.objects = ptr(arrays.objects).value
    g.object_count = len(arrays.objects)
    g.leaf_material = ptr(arrays.leaf_material).value
    for i in range(3):
        g.lower[i] = float(lower[i])
        g.size[i] = float(size[i])
    R.materials = C.cast(marr, P).value
    R.material_count = len(mats)
    R.lights = C.cast(larr, P).value
    R.light_count = len(lights)
    R.camera = cam
    R.sky_zenith[:] = rs.sky_zenith[:]
    R.sky_horizon[:] = rs.sky_horizon[:]
    R.sky_ground[:] = rs.sky_ground[:]
    rgb = np.zeros((y1 - y0, x1 - x0, 3), np.float64)
    L.or_render64_sequenced(C.byref(R), width, height, x0, y0, x1, y1, first, spp, seed, surfaces, min_depth, 1 if sobol else 0,
                            ptr(rgb), threads or min(16, os.cpu_count() or 1))
    return rgb
