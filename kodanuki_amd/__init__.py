"""kodanuki_amd — the MI355X render hot path behind Kodanuki's plugin hooks.

This module is the Python face of libkodanuki_hip.so (include/kodanuki.h):
ctypes mirrors of the C structs and thin wrappers that raise on a nonzero
return code. It loads the in-tree library and fails loudly when it is not
built; there is no CPU fallback.

The objects mirror the reference's vocabulary: a Scene is what loadScene
leaves behind (source/syntax/evaluator.h:369) with its object programs
compiled (source/sdf/scene.cpp:290-296); a Device holds the scene, its
octree and the integrator state in HBM, and answers extend (rays → hits)
and pixel passes (source/core/plugin.h:168-357).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkodanuki_hip.so")

KDN_HIT_MISSED, KDN_HIT_SURFACE, KDN_HIT_SOLID = 0, 1, 2
NO_LEAF = 0xFFFFFFFF
KIND_EMPTY, KIND_FILLED, KIND_SURFACE, KIND_BRANCH = 0, 1, 2, 3


class Step(C.Structure):
    _fields_ = [("op", C.c_uint8), ("flags", C.c_uint8), ("a", C.c_uint16), ("b", C.c_uint16),
                ("c", C.c_uint16), ("value", C.c_float), ("leaf", C.c_uint32)]


class Object(C.Structure):
    _fields_ = [("first", C.c_uint32), ("count", C.c_uint32), ("result", C.c_uint32), ("material", C.c_uint32)]


class Ray(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("direction", C.c_float * 3), ("far_", C.c_float),
                ("inside", C.c_uint32), ("within", C.c_uint32)]


class Hit(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("distance", C.c_float), ("point", C.c_float * 3), ("normal", C.c_float * 3),
                ("leaf", C.c_uint32), ("other", C.c_uint32), ("share", C.c_float), ("cell", C.c_uint32)]


class Material(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("albedo", C.c_float * 3), ("roughness", C.c_float), ("refraction", C.c_float),
                ("dispersion", C.c_float), ("glow", C.c_float * 3), ("strength", C.c_float)]


class Light(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("at", C.c_float * 3), ("along", C.c_float * 3), ("across", C.c_float * 3),
                ("power", C.c_float * 3), ("direction", C.c_float * 3), ("radius", C.c_float), ("angle", C.c_float)]


LIGHT_POINT, LIGHT_QUAD, LIGHT_DIRECTIONAL, LIGHT_SPHERE, LIGHT_SUN, LIGHT_MOON = 0, 1, 2, 3, 4, 5


class Leaf(C.Structure):
    _fields_ = [("primitive", C.c_uint32), ("material", C.c_uint32), ("args", C.c_double * 8), ("linear", C.c_double * 9),
                ("origin", C.c_double * 3), ("stretch", C.c_double)]


class Node(C.Structure):
    _fields_ = [("join", C.c_uint32), ("radius", C.c_double), ("leaf", C.c_uint32), ("first_child", C.c_uint32),
                ("child_count", C.c_uint32)]


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3), ("fov", C.c_float)]


class RenderSettings(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("seed", C.c_int32), ("samples", C.c_int32),
                ("surfaces", C.c_int32), ("far_", C.c_float), ("sky_zenith", C.c_float * 3),
                ("sky_horizon", C.c_float * 3), ("sky_ground", C.c_float * 3), ("batch", C.c_int32), ("exposure", C.c_float),
                ("curve", C.c_int32), ("tolerance", C.c_float), ("split", C.c_int32)]


CURVE_FILMIC, CURVE_CLAMPED = 0, 1
MATTE, PLASTIC, METAL, GLASS = 0, 1, 2, 3  # kdn_material.kind


class SceneInfo(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("objects", C.c_uint32), ("leaves", C.c_uint32), ("nodes", C.c_uint32),
                ("top_nodes", C.c_uint32), ("children", C.c_uint32), ("materials", C.c_uint32), ("lights", C.c_uint32)]


class OctreeStats(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("nodes", C.c_uint32 * 16), ("surface", C.c_uint32 * 16),
                ("filled", C.c_uint32 * 16), ("entries", C.c_uint32), ("pool_steps", C.c_uint64), ("build_ms", C.c_double)]


class NodeRecord(C.Structure):
    _fields_ = [("level", C.c_uint32), ("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32), ("kind", C.c_uint32),
                ("first_child", C.c_uint32), ("entry_first", C.c_uint32), ("entry_count", C.c_uint32)]


class EntryRecord(C.Structure):
    _fields_ = [("object", C.c_uint32), ("program", C.c_uint32)]


class PassDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("first_sample", C.c_uint32), ("spp", C.c_uint32),
                ("seed", C.c_uint32), ("surfaces", C.c_uint32), ("min_depth", C.c_uint32), ("flags", C.c_uint32),
                ("batch", C.c_uint32), ("round_paths", C.c_uint32)]


# kdn_estimate: a pixel's estimate as the reference's frame keeps it
ESTIMATE_DTYPE = np.dtype([("total", "<f8", 3), ("taken", "<i4"), ("batches", "<i4"), ("batch", "<f8", 3), ("means", "<f8", 3),
                           ("squares", "<f8", 3), ("shared", "<f8"), ("shared_means", "<f8"), ("shared_squares", "<f8"),
                           ("own_squares", "<f8")])
assert ESTIMATE_DTYPE.itemsize == 136


class PassStats(C.Structure):
    _fields_ = [("ms", C.c_double), ("segments", C.c_uint64), ("shadow_rays", C.c_uint64), ("bounces", C.c_uint32),
                ("trace_launches", C.c_uint32), ("trace_ms", C.c_double), ("shadow_ms", C.c_double), ("shade_ms", C.c_double),
                ("other_ms", C.c_double), ("trace_flops", C.c_double), ("shadow_flops", C.c_double),
                ("finish_ms", C.c_double), ("finish_flops", C.c_double), ("shadow_launches", C.c_uint32),
                ("finish_launches", C.c_uint32), ("guide_ms", C.c_double), ("guide_infer_ms", C.c_double),
                ("guide_loss", C.c_double), ("guide_queries", C.c_uint64), ("guide_samples", C.c_uint32),
                ("guide_steps", C.c_uint32)]


PASS_KEEP, PASS_TIME, PASS_COUNT, PASS_NO_FINISH, PASS_GUIDE, PASS_HIGHLIGHT, PASS_ESTIMATES = 1, 2, 4, 8, 16, 32, 64
PASS_SOBOL, PASS_COUNTED, PASS_RECORD = 128, 256, 512


class Stopping(C.Structure):
    """kdn_stopping: the confidence-interval member's settings (the
    reference's Stopping, core/stopping.h:102-132), split 1."""
    _fields_ = [("samples", C.c_int32), ("batch", C.c_int32), ("tolerance", C.c_double), ("exposure", C.c_float),
                ("curve", C.c_int32), ("split", C.c_int32)]


def stopping_effort(tolerance):
    """effortFor: (samples, batch) a tolerance asks for where the scene names none."""
    samples, batch = C.c_int32(), C.c_int32()
    _check(library().kdn_stopping_effort(C.c_double(tolerance), C.byref(samples), C.byref(batch)))
    return samples.value, batch.value


def stopping_of(settings, split=1):
    """The Stopping a scene's render settings ask for (frameStopping,
    core/frame.cpp:304-317): effortFor's ceiling and batch where the scene
    names none; every sample here draws its own first vertex (split 1)."""
    samples, batch = stopping_effort(settings.tolerance)
    if settings.batch > 0:
        batch = settings.batch
    if settings.samples > 0:
        samples = settings.samples
    if split > 0 and batch % split:
        raise KodanukiError(f"a split of {split} needs batches of a multiple of it, not {batch}")
    return Stopping(samples, batch, settings.tolerance, settings.exposure, settings.curve, split)


def stopping_trial(stopping):
    """confidenceTrial: samples each pixel's trial draws (0: none)."""
    n = C.c_int32()
    _check(library().kdn_stopping_trial(C.byref(stopping), C.byref(n)))
    return n.value


class GuideInfo(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("features", C.c_uint32), ("table_size", C.c_uint32), ("width", C.c_uint32),
                ("layers", C.c_uint32), ("lobes", C.c_uint32), ("depths", C.c_uint32), ("train_paths", C.c_uint32),
                ("resolution", C.c_uint32 * 16), ("dense_levels", C.c_uint32), ("steps", C.c_uint32),
                ("loss_scale", C.c_float), ("share", C.c_float), ("last_loss", C.c_double), ("last_samples", C.c_uint32),
                ("pad", C.c_uint32), ("grid_params", C.c_uint64), ("mlp_params", C.c_uint64), ("lower", C.c_float * 3),
                ("size", C.c_float * 3)]


DENOISE_PIXEL_DTYPE = np.dtype([("colour", "<f8", 3), ("moment", "<f8"), ("length", "<f4"), ("distance", "<f4"),
                                ("normal", "<f8", 3), ("leaf", "<u4"), ("pad", "<u4"), ("albedo", "<f8", 3)])
assert DENOISE_PIXEL_DTYPE.itemsize == 96


class DenoiseFrame(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("forward", C.c_double * 3), ("right", C.c_double * 3),
                ("up", C.c_double * 3), ("reach", C.c_double), ("aspect", C.c_double), ("centre_x", C.c_double),
                ("centre_y", C.c_double), ("demodulated", C.c_uint32), ("pad", C.c_uint32)]


class DenoiseStats(C.Structure):
    _fields_ = [("gbuffer_ms", C.c_double), ("filter_ms", C.c_double)]


# kdn_surface: what a camera ray met first (the reference's Surface)
SURFACE_DTYPE = np.dtype([("distance", "<f4"), ("normal", "<f4", 3), ("leaf", "<u4"), ("albedo", "<f4", 3)])
assert SURFACE_DTYPE.itemsize == 32


class Lens(C.Structure):
    _fields_ = [("eye", C.c_double * 3), ("forward", C.c_double * 3), ("right", C.c_double * 3), ("up", C.c_double * 3),
                ("reach", C.c_double), ("aspect", C.c_double)]


def lens_of(eye, target, up, fov, width, height):
    """The pinhole lens of a camera at an image's size (lensOf, core/camera.cpp:9-25)."""
    eye = np.asarray(eye, np.float64)
    forward = np.asarray(target, np.float64) - eye
    forward = forward / np.sqrt(forward @ forward)
    up = np.asarray(up, np.float64)
    right = np.array([forward[1] * up[2] - forward[2] * up[1], forward[2] * up[0] - forward[0] * up[2],
                      forward[0] * up[1] - forward[1] * up[0]])
    right = right / np.sqrt(right @ right)
    upv = np.array([right[1] * forward[2] - right[2] * forward[1], right[2] * forward[0] - right[0] * forward[2],
                    right[0] * forward[1] - right[1] * forward[0]])
    lens = Lens()
    lens.eye[:], lens.forward[:], lens.right[:], lens.up[:] = list(eye), list(forward), list(right), list(upv)
    lens.reach = float(np.tan(np.float64(np.float32(fov)) / 2))
    lens.aspect = width / height
    return lens


class TemporalDesc(C.Structure):
    _fields_ = [("moved", C.c_uint32), ("restarted", C.c_uint32), ("recounted", C.c_uint32), ("denoise", C.c_uint32)]


class TemporalStats(C.Structure):
    _fields_ = [("gbuffer_ms", C.c_double), ("temporal_ms", C.c_double), ("filter_ms", C.c_double)]


# a pixel's packed plain guide: count, material, share, unused, then 12
# lumps of axis (3), across (3), point (3), mean (2), factor (3), weight
HG_FLOATS, HG_HEAD, HG_LUMP_FLOATS, HG_LUMPS = 184, 4, 15, 12


class HighlightStats(C.Structure):
    _fields_ = [("ms", C.c_double), ("survey_points", C.c_uint32), ("highlights", C.c_uint32),
                ("guided_pixels", C.c_uint32), ("pad", C.c_uint32), ("mean_lumps", C.c_double), ("mean_share", C.c_double)]


class GuideConfig(C.Structure):
    _fields_ = [("learning_rate", C.c_float), ("share", C.c_float), ("epochs", C.c_float), ("min_steps", C.c_uint32),
                ("max_steps", C.c_uint32), ("max_resolution", C.c_uint32), ("train_paths", C.c_uint32),
                ("learn_select", C.c_uint32), ("product", C.c_uint32), ("pad", C.c_uint32)]


class GuideQuery(C.Structure):
    _fields_ = [("point", C.c_float * 3), ("out", C.c_float * 3), ("normal", C.c_float * 3), ("nm", C.c_float)]


class GuideSample(C.Structure):
    _fields_ = [("point", C.c_float * 3), ("out", C.c_float * 3), ("normal", C.c_float * 3), ("nm", C.c_float),
                ("in_", C.c_float * 3), ("weight", C.c_float), ("bsdf_pdf", C.c_float), ("pad", C.c_float)]


class GuideLobe(C.Structure):
    _fields_ = [("axis", C.c_float * 3), ("sharpness", C.c_float), ("share", C.c_float)]


class JobDesc(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_uint32), ("seed", C.c_uint32),
                ("surfaces", C.c_uint32), ("min_depth", C.c_uint32), ("depth", C.c_int32), ("flags", C.c_uint32),
                ("pad", C.c_uint32)]


JOB_PLAIN_GUIDE = 1
JOB_LEARNED_GUIDE = 2


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_uint32), ("samples_done", C.c_uint32), ("submitted_ms", C.c_double),
                ("admitted_ms", C.c_double), ("finished_ms", C.c_double)]


class ServeStats(C.Structure):
    _fields_ = [("pass_", PassStats), ("paths", C.c_uint64), ("resident_bytes", C.c_uint64), ("rounds", C.c_uint32),
                ("jobs_in_round", C.c_uint32), ("slices_in_round", C.c_uint32), ("waiting", C.c_uint32),
                ("running", C.c_uint32), ("pad", C.c_uint32), ("ms", C.c_double)]


JOB_WAITING, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, 3


# numpy views of the structs
STEP_DTYPE = np.dtype([("op", "u1"), ("flags", "u1"), ("a", "<u2"), ("b", "<u2"), ("c", "<u2"), ("value", "<f4"), ("leaf", "<u4")])
RAY_DTYPE = np.dtype([("origin", "<f4", 3), ("direction", "<f4", 3), ("far_", "<f4"), ("inside", "<u4"), ("within", "<u4")])
HIT_DTYPE = np.dtype([("kind", "<u4"), ("distance", "<f4"), ("point", "<f4", 3), ("normal", "<f4", 3), ("leaf", "<u4"),
                      ("other", "<u4"), ("share", "<f4"), ("cell", "<u4")])
NODE_DTYPE = np.dtype([("level", "<u4"), ("x", "<u4"), ("y", "<u4"), ("z", "<u4"), ("kind", "<u4"), ("first_child", "<u4"),
                       ("entry_first", "<u4"), ("entry_count", "<u4")])
ENTRY_DTYPE = np.dtype([("object", "<u4"), ("program", "<u4")])
OBJECT_DTYPE = np.dtype([("first", "<u4"), ("count", "<u4"), ("result", "<u4"), ("material", "<u4")])
# kdn_chain_query and kdn_chain
CHAIN_QUERY_DTYPE = np.dtype([("from", "<f8", 3), ("sideways", "<f8", (2, 3)), ("seed", "<f8", 3), ("width", "<f8"),
                              ("leaf", "<u4"), ("light", "<u4")])
CHAIN_DTYPE = np.dtype([("found", "<u4"), ("pad", "<u4"), ("point", "<f8", 3), ("axis", "<f8", 3), ("across", "<f8", 3),
                        ("up", "<f8", 3), ("spread", "<f8", 3), ("motion", "<f8", (2, 3))])
assert STEP_DTYPE.itemsize == 16 and RAY_DTYPE.itemsize == 36 and HIT_DTYPE.itemsize == 48
assert CHAIN_QUERY_DTYPE.itemsize == 112 and CHAIN_DTYPE.itemsize == 176

# Every symbol include/kodanuki.h declares.
EXPORTS = [
    "kdn_scene_load", "kdn_scene_free", "kdn_scene_describe", "kdn_scene_tape", "kdn_scene_tape_values",
    "kdn_scene_leaves", "kdn_scene_leaf_program", "kdn_scene_nodes", "kdn_scene_materials", "kdn_scene_lights",
    "kdn_scene_settings", "kdn_scene_region", "kdn_open", "kdn_close", "kdn_last_error", "kdn_global_error", "kdn_upload_scene",
    "kdn_build_octree", "kdn_octree_sizes", "kdn_octree_download", "kdn_extend", "kdn_extend_device",
    "kdn_render_pass", "kdn_read_accumulation", "kdn_read_image", "kdn_read_estimates", "kdn_fold_samples",
    "kdn_stopping_effort", "kdn_stopping_trial", "kdn_stopping_allot", "kdn_read_stopping", "kdn_sequence_numbers",
    "kdn_image_write_ppm", "kdn_image_encode", "kdn_image_write_pfm",
    "kdn_denoise", "kdn_denoise_device", "kdn_denoise_device_static", "kdn_denoise_accumulation",
    "kdn_set_camera", "kdn_temporal_frame", "kdn_temporal_history", "kdn_temporal_step",
    "kdn_guide_reset", "kdn_guide_configure", "kdn_guide_describe", "kdn_guide_params", "kdn_guide_set_params", "kdn_guide_infer",
    "kdn_guide_gradients", "kdn_guide_train",
    "kdn_serve_open", "kdn_serve_submit", "kdn_serve_step", "kdn_serve_status", "kdn_serve_result",
    "kdn_serve_release", "kdn_serve_error", "kdn_serve_job_error", "kdn_serve_close",
    "kdn_highlight_guides", "kdn_highlight_guide_at", "kdn_highlight_density", "kdn_highlight_draw", "kdn_highlight_chains",
]

_lib = None


class KodanukiError(RuntimeError):
    pass


def library():
    """Load libkodanuki_hip.so from the package folder (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise KodanukiError(f"{LIB_PATH} is not built: run __graft_entry__.build() (hipcc, gfx950)")
    return _bind(C.CDLL(LIB_PATH))


def _bind(lib):
    """Declare the C ABI's signatures on a loaded library and keep it."""
    global _lib
    P = C.c_void_p
    sig = {
        "kdn_scene_load": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(P), C.c_char_p, C.c_size_t]),
        "kdn_scene_free": (None, [P]),
        "kdn_scene_describe": (C.c_int, [P, C.POINTER(SceneInfo)]),
        "kdn_scene_tape": (C.c_int, [P, P, P]),
        "kdn_scene_tape_values": (C.c_int, [P, P]),
        "kdn_scene_leaves": (C.c_int, [P, P]),
        "kdn_scene_leaf_program": (C.c_int, [P, C.c_uint32, P, P, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
        "kdn_scene_nodes": (C.c_int, [P, P, P, P]),
        "kdn_scene_materials": (C.c_int, [P, P]),
        "kdn_scene_lights": (C.c_int, [P, P]),
        "kdn_scene_settings": (C.c_int, [P, C.POINTER(Camera), C.POINTER(RenderSettings)]),
        "kdn_scene_region": (C.c_int, [P, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "kdn_open": (C.c_int, [C.c_int, C.POINTER(P)]),
        "kdn_close": (None, [P]),
        "kdn_last_error": (C.c_int, [P, C.c_char_p, C.c_size_t]),
        "kdn_global_error": (C.c_char_p, []),
        "kdn_upload_scene": (C.c_int, [P, P]),
        "kdn_build_octree": (C.c_int, [P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, C.POINTER(OctreeStats)]),
        "kdn_octree_sizes": (C.c_int, [P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "kdn_octree_download": (C.c_int, [P, P, P, P]),
        "kdn_extend": (C.c_int, [P, P, C.c_uint32, P]),
        "kdn_extend_device": (C.c_int, [P, P, C.c_uint32, P, P]),
        "kdn_render_pass": (C.c_int, [P, C.POINTER(PassDesc), C.POINTER(PassStats)]),
        "kdn_read_accumulation": (C.c_int, [P, P, P, P]),
        "kdn_read_image": (C.c_int, [P, P]),
        "kdn_read_estimates": (C.c_int, [P, P]),
        "kdn_fold_samples": (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.c_uint32, P]),
        "kdn_stopping_effort": (C.c_int, [C.c_double, P, P]),
        "kdn_stopping_trial": (C.c_int, [P, P]),
        "kdn_stopping_allot": (C.c_int, [P, P, C.c_int32, P]),
        "kdn_read_stopping": (C.c_int, [P, P, P, P]),
        "kdn_sequence_numbers": (C.c_int, [P, C.c_uint64, P, C.c_uint32, C.c_uint32, C.c_int32, P]),
        "kdn_image_write_ppm": (C.c_int, [P, C.c_int32, C.c_int32, C.c_float, C.c_int32, C.c_char_p]),
        "kdn_image_encode": (C.c_int, [P, C.c_int32, C.c_int32, C.c_float, C.c_int32, P]),
        "kdn_image_write_pfm": (C.c_int, [P, C.c_int32, C.c_int32, C.c_char_p]),
        "kdn_denoise": (C.c_int, [P, C.POINTER(DenoiseFrame), P, P]),
        "kdn_denoise_device": (C.c_int, [P, C.POINTER(DenoiseFrame), P, P, P]),
        "kdn_denoise_device_static": (C.c_int, [P, C.POINTER(DenoiseFrame), P, P, P]),
        "kdn_denoise_accumulation": (C.c_int, [P, P, C.POINTER(DenoiseStats)]),
        "kdn_set_camera": (C.c_int, [P, P]),
        "kdn_temporal_frame": (C.c_int, [P, P, P, P, P]),
        "kdn_temporal_history": (C.c_int, [P, P]),
        "kdn_temporal_step": (C.c_int, [P, C.c_int32, C.c_int32, P, P, P, P, P, P, P, C.c_int32, C.c_int32, P, P]),
        "kdn_highlight_guides": (C.c_int, [P, C.c_int32, C.c_int32, C.POINTER(HighlightStats)]),
        "kdn_highlight_guide_at": (C.c_int, [P, C.c_int32, C.c_int32, P]),
        "kdn_highlight_density": (C.c_int, [P, P, P, P, P, P, C.c_uint32, P]),
        "kdn_highlight_draw": (C.c_int, [P, P, P, P, C.c_uint32, P]),
        "kdn_highlight_chains": (C.c_int, [P, P, C.c_uint32, P]),
        "kdn_guide_reset": (C.c_int, [P, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "kdn_guide_describe": (C.c_int, [P, C.POINTER(GuideInfo)]),
        "kdn_guide_configure": (C.c_int, [P, C.POINTER(GuideConfig)]),
        "kdn_guide_params": (C.c_int, [P, P, P]),
        "kdn_guide_set_params": (C.c_int, [P, P, P]),
        "kdn_guide_infer": (C.c_int, [P, P, C.c_uint32, P, P]),
        "kdn_guide_gradients": (C.c_int, [P, P, C.c_uint32, P, P, C.POINTER(C.c_double)]),
        "kdn_guide_train": (C.c_int, [P, P, C.c_uint32, C.c_uint32, C.c_uint32, P]),
        "kdn_serve_open": (C.c_int, [P, C.c_uint64, C.c_uint32, C.POINTER(P)]),
        "kdn_serve_submit": (C.c_int, [P, C.POINTER(JobDesc), C.POINTER(C.c_uint64)]),
        "kdn_serve_step": (C.c_int, [P, C.POINTER(ServeStats)]),
        "kdn_serve_status": (C.c_int, [P, C.c_uint64, C.POINTER(JobStatus)]),
        "kdn_serve_result": (C.c_int, [P, C.c_uint64, P, P, P]),
        "kdn_serve_release": (C.c_int, [P, C.c_uint64]),
        "kdn_serve_error": (C.c_char_p, [P]),
        "kdn_serve_job_error": (C.c_char_p, [P, C.c_uint64]),
        "kdn_serve_close": (None, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc, ctx=None):
    if rc != 0:
        lib = library()
        if ctx is not None:
            buf = C.create_string_buffer(1024)
            lib.kdn_last_error(ctx, buf, 1024)
            msg = buf.value.decode()
        else:
            msg = lib.kdn_global_error().decode()
        raise KodanukiError(f"kodanuki error {rc}: {msg}")


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Scene:
    """A loaded and compiled .sdf scene (host memory)."""

    def __init__(self, text, base_dir=None):
        lib = library()
        self._h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = lib.kdn_scene_load(text.encode(), base_dir.encode() if base_dir else None, C.byref(self._h), err, 1024)
        if rc != 0:
            raise KodanukiError(f"scene does not load: {err.value.decode()}")
        self.info = SceneInfo()
        _check(lib.kdn_scene_describe(self._h, C.byref(self.info)))

    @classmethod
    def from_file(cls, path):
        with open(path) as f:
            return cls(f.read(), os.path.dirname(os.path.abspath(path)))

    def __del__(self):
        if getattr(self, "_h", None):
            library().kdn_scene_free(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def tape(self):
        """(steps, objects) as numpy structured arrays."""
        steps = np.zeros(self.info.steps, STEP_DTYPE)
        objects = np.zeros(self.info.objects, OBJECT_DTYPE)
        _check(library().kdn_scene_tape(self._h, _ptr(steps), _ptr(objects)))
        return steps, objects

    def tape_values(self):
        v = np.zeros(self.info.steps, np.float64)
        _check(library().kdn_scene_tape_values(self._h, _ptr(v)))
        return v

    def leaves(self):
        arr = (Leaf * max(1, self.info.leaves))()
        _check(library().kdn_scene_leaves(self._h, C.cast(arr, C.c_void_p)))
        return list(arr)[: self.info.leaves]

    def leaf_program(self, leaf):
        """One leaf alone over world space (the survey's chains walk it):
        its steps (STEP_DTYPE), their constants in double, its result."""
        count, result = C.c_uint32(), C.c_uint32()
        _check(library().kdn_scene_leaf_program(self._h, leaf, None, None, 0, C.byref(count), C.byref(result)))
        steps = np.zeros(max(1, count.value), STEP_DTYPE)
        values = np.zeros(max(1, count.value), np.float64)
        _check(library().kdn_scene_leaf_program(self._h, leaf, _ptr(steps), _ptr(values), count.value, C.byref(count),
                                                C.byref(result)))
        return steps[: count.value], values[: count.value], result.value

    def nodes(self):
        nodes = (Node * max(1, self.info.nodes))()
        tops = np.zeros(max(1, self.info.top_nodes), np.uint32)
        children = np.zeros(max(1, self.info.children), np.uint32)
        _check(library().kdn_scene_nodes(self._h, C.cast(nodes, C.c_void_p), _ptr(tops), _ptr(children)))
        return list(nodes)[: self.info.nodes], tops[: self.info.top_nodes], children[: self.info.children]

    def materials(self):
        arr = (Material * max(1, self.info.materials))()
        if self.info.materials:
            _check(library().kdn_scene_materials(self._h, C.cast(arr, C.c_void_p)))
        return list(arr)[: self.info.materials]

    def lights(self):
        arr = (Light * max(1, self.info.lights))()
        if self.info.lights:
            _check(library().kdn_scene_lights(self._h, C.cast(arr, C.c_void_p)))
        return list(arr)[: self.info.lights]

    def settings(self):
        cam, rs = Camera(), RenderSettings()
        _check(library().kdn_scene_settings(self._h, C.byref(cam), C.byref(rs)))
        return cam, rs

    def region(self, depth):
        """(lower[3], size[3]) of the octree's box for a depth cap."""
        lower = (C.c_float * 3)()
        size = (C.c_float * 3)()
        _check(library().kdn_scene_region(self._h, depth, lower, size))
        return np.array(lower[:], np.float32), np.array(size[:], np.float32)


def hip_runtimes():
    """The HIP runtime libraries mapped into this process."""
    try:
        with open("/proc/self/maps") as f:
            maps = f.read()
    except OSError:
        return []
    return sorted({line.split()[-1] for line in maps.splitlines() if "libamdhip64" in line and line.split()[-1].startswith("/")})


class Device:
    """A context on one HIP device: scene, octree and integrator in HBM."""

    def __init__(self, device=0):
        lib = library()
        runtimes = hip_runtimes()
        if len(runtimes) > 1:
            # two HIP runtimes (e.g. one a Python package bundles, loaded
            # after this library) would manage the same device apart
            raise KodanukiError("two HIP runtimes are loaded in this process (" + ", ".join(runtimes) +
                                "): open the device before importing a package that bundles its own")
        self._h = C.c_void_p()
        _check(lib.kdn_open(device, C.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None):
            library().kdn_close(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def handle(self):
        return self._h

    def upload(self, scene):
        _check(library().kdn_upload_scene(self._h, scene.handle), self._h)

    def build_octree(self, lower, size, depth):
        stats = OctreeStats()
        lo = (C.c_float * 3)(*[float(x) for x in lower])
        sz = (C.c_float * 3)(*[float(x) for x in size])
        _check(library().kdn_build_octree(self._h, lo, sz, depth, C.byref(stats)), self._h)
        return stats

    def octree(self):
        """(nodes, entries, pool) as numpy structured arrays."""
        n, e, p = C.c_uint64(), C.c_uint64(), C.c_uint64()
        lib = library()
        _check(lib.kdn_octree_sizes(self._h, C.byref(n), C.byref(e), C.byref(p)), self._h)
        nodes = np.zeros(n.value, NODE_DTYPE)
        entries = np.zeros(e.value, ENTRY_DTYPE)
        pool = np.zeros(p.value, STEP_DTYPE)
        _check(lib.kdn_octree_download(self._h, _ptr(nodes), _ptr(entries), _ptr(pool)), self._h)
        return nodes, entries, pool

    def extend(self, rays):
        rays = np.ascontiguousarray(rays, RAY_DTYPE)
        hits = np.zeros(len(rays), HIT_DTYPE)
        _check(library().kdn_extend(self._h, _ptr(rays), len(rays), _ptr(hits)), self._h)
        return hits

    def extend_device(self, rays_ptr, count, hits_ptr, stream=None):
        _check(library().kdn_extend_device(self._h, C.c_void_p(rays_ptr), count, C.c_void_p(hits_ptr),
                                           C.c_void_p(stream) if stream else None), self._h)

    def render_pass(self, width, height, spp, first_sample=0, seed=1, surfaces=7, min_depth=4, keep=False, flags=0, batch=0,
                    round_paths=0):
        self._image_size = (width, height)
        d = PassDesc(width, height, first_sample, spp, seed, surfaces, min_depth, (PASS_KEEP if keep else 0) | flags, batch,
                     round_paths)
        st = PassStats()
        _check(library().kdn_render_pass(self._h, C.byref(d), C.byref(st)), self._h)
        return st

    def estimates(self, width, height):
        """The per-pixel estimates passes with PASS_ESTIMATES kept: [height, width] of ESTIMATE_DTYPE."""
        out = np.zeros((height, width), ESTIMATE_DTYPE)
        _check(library().kdn_read_estimates(self._h, _ptr(out)), self._h)
        return out

    def fold_samples(self, colours, batch, estimates=None):
        """colours [pixels, spp, 3] (float64) folded in sample order into
        estimates [pixels] (ESTIMATE_DTYPE; fresh when None) on the device."""
        col = np.ascontiguousarray(colours, np.float64)
        pixels, spp, _ = col.shape
        est = np.zeros(pixels, ESTIMATE_DTYPE) if estimates is None else np.ascontiguousarray(estimates, ESTIMATE_DTYPE).copy()
        _check(library().kdn_fold_samples(self._h, _ptr(col), pixels, spp, batch, _ptr(est)), self._h)
        return est

    # ---- adaptive stopping and the sequence ----
    def stopping_allot(self, stopping, trial):
        """Judge the trial estimates and allot every pixel's count on the
        device; returns the count pass's paths."""
        total = C.c_uint64()
        _check(library().kdn_stopping_allot(self._h, C.byref(stopping), trial, C.byref(total)), self._h)
        return total.value

    def stopping_state(self, width, height):
        """The last allotment: (wanted [h, w] f32, spreads [h, w, 2] f32, allotted [h, w] i32)."""
        wanted = np.zeros((height, width), np.float32)
        spreads = np.zeros((height, width, 2), np.float32)
        allotted = np.zeros((height, width), np.int32)
        _check(library().kdn_read_stopping(self._h, _ptr(wanted), _ptr(spreads), _ptr(allotted)), self._h)
        return wanted, spreads, allotted

    def render_stopped(self, width, height, stopping, seed=1, surfaces=7, min_depth=4, flags=0, round_paths=0):
        """A still frame as the reference's frame draws it with the
        confidence-interval member (core/frame.cpp:28-62, 435): a trial of two
        batches in every pixel, then each pixel's allotment less its trial;
        one pass of stopping.samples where the member draws no trial.
        Returns (trial, paths of the count pass, [pass stats])."""
        trial = stopping_trial(stopping)
        if trial <= 0:
            return 0, 0, [self.render_pass(width, height, stopping.samples, 0, seed, surfaces, min_depth, flags=flags | PASS_ESTIMATES,
                                           batch=stopping.batch)]
        first = self.render_pass(width, height, trial, 0, seed, surfaces, min_depth, flags=flags | PASS_ESTIMATES, batch=stopping.batch)
        total = self.stopping_allot(stopping, trial)
        rest = self.render_pass(width, height, 0, trial, seed, surfaces, min_depth, keep=True,
                                flags=flags | PASS_ESTIMATES | PASS_COUNTED, batch=stopping.batch, round_paths=round_paths)
        return trial, total, [first, rest]

    def sequence_numbers(self, stream, indices, coordinate, sobol=True):
        """numberAt of an eye path on the device: float64 per sample index."""
        idx = np.ascontiguousarray(indices, np.uint32)
        out = np.zeros(len(idx), np.float64)
        _check(library().kdn_sequence_numbers(self._h, C.c_uint64(stream), _ptr(idx), len(idx), coordinate, 1 if sobol else 0,
                                              _ptr(out)), self._h)
        return out

    def accumulation(self, width, height):
        rgb = np.zeros((height, width, 3), np.float64)
        sq = np.zeros((height, width, 3), np.float64)
        n = np.zeros((height, width), np.uint32)
        _check(library().kdn_read_accumulation(self._h, _ptr(rgb), _ptr(sq), _ptr(n)), self._h)
        return rgb, sq, n

    # ---- the denoiser ----
    def denoise(self, pixels, forward, right, up, reach, centre=(0.5, 0.5), demodulated=False):
        """pixels [height, width] of DENOISE_PIXEL_DTYPE -> rgb [height, width, 3]."""
        px = np.ascontiguousarray(pixels, DENOISE_PIXEL_DTYPE)
        h, w = px.shape
        f = DenoiseFrame(w, h, (C.c_double * 3)(*forward), (C.c_double * 3)(*right), (C.c_double * 3)(*up), reach, w / h,
                         centre[0], centre[1], 1 if demodulated else 0, 0)
        rgb = np.zeros((h, w, 3), np.float32)
        _check(library().kdn_denoise(self._h, C.byref(f), _ptr(px), _ptr(rgb)), self._h)
        return rgb

    def denoise_device(self, width, height, pixels_ptr, rgb_ptr, forward, right, up, reach, centre=(0.5, 0.5),
                       demodulated=False, static=False, stream=None):
        """The filter on device buffers (pointers): width x height history
        pixels in, rgb out. static: the entry point that allocates nothing
        (frames up to 2^20 pixels)."""
        f = DenoiseFrame(width, height, (C.c_double * 3)(*forward), (C.c_double * 3)(*right), (C.c_double * 3)(*up), reach,
                         width / height, centre[0], centre[1], 1 if demodulated else 0, 0)
        call = library().kdn_denoise_device_static if static else library().kdn_denoise_device
        _check(call(self._h, C.byref(f), C.c_void_p(pixels_ptr), C.c_void_p(rgb_ptr), C.c_void_p(stream) if stream else None),
               self._h)

    def denoise_accumulation(self, width, height, stats=False):
        rgb = np.zeros((height, width, 3), np.float32)
        st = DenoiseStats()
        _check(library().kdn_denoise_accumulation(self._h, _ptr(rgb), C.byref(st) if stats else None), self._h)
        return (rgb, st) if stats else rgb

    # ---- the temporal member (a viewer's frames) ----
    def set_camera(self, eye, target, up, fov):
        cam = Camera()
        cam.eye[:], cam.target[:], cam.up[:] = list(eye), list(target), list(up)
        cam.fov = fov
        _check(library().kdn_set_camera(self._h, C.byref(cam)), self._h)

    def temporal_frame(self, moved=False, restarted=False, recounted=False, denoise=False, stats=False):
        """One viewer frame after render_pass(..., flags=PASS_RECORD): the
        shown image [h, w, 3] (and the denoised one with denoise)."""
        w, h = self._image_size
        shown = np.zeros((h, w, 3), np.float32)
        denoised = np.zeros((h, w, 3), np.float32) if denoise else None
        d = TemporalDesc(int(moved), int(restarted), int(recounted), int(denoise))
        st = TemporalStats()
        _check(library().kdn_temporal_frame(self._h, C.byref(d), _ptr(shown), _ptr(denoised) if denoise else None,
                                            C.byref(st) if stats else None), self._h)
        out = (shown, denoised) if denoise else shown
        return (out, st) if stats else out

    def temporal_history(self, width, height):
        px = np.zeros((height, width), DENOISE_PIXEL_DTYPE)
        _check(library().kdn_temporal_history(self._h, _ptr(px)), self._h)
        return px

    def temporal_step(self, lens, before_lens, parts, record, gbuffer, before=None, moved=False, restarted=False,
                      recounted=False):
        """The temporal step on given frames: parts [h, w, 3] f64, record and
        gbuffer [h, w] SURFACE_DTYPE, before None or [bh, bw]
        DENOISE_PIXEL_DTYPE -> (history [h, w], shown [h, w, 3])."""
        parts = np.ascontiguousarray(parts, np.float64)
        h, w, _ = parts.shape
        record = np.ascontiguousarray(record, SURFACE_DTYPE)
        gbuffer = np.ascontiguousarray(gbuffer, SURFACE_DTYPE)
        bp = None if before is None else np.ascontiguousarray(before, DENOISE_PIXEL_DTYPE)
        out = np.zeros((h, w), DENOISE_PIXEL_DTYPE)
        shown = np.zeros((h, w, 3), np.float32)
        d = TemporalDesc(int(moved), int(restarted), int(recounted), 0)
        _check(library().kdn_temporal_step(self._h, w, h, C.byref(lens), C.byref(before_lens), C.byref(d), _ptr(parts), _ptr(record),
                                           _ptr(gbuffer), _ptr(bp) if bp is not None else None,
                                           bp.shape[1] if bp is not None else 0, bp.shape[0] if bp is not None else 0,
                                           _ptr(out), _ptr(shown)), self._h)
        return out, shown

    # ---- the plain (highlight) guide ----
    def highlight_guides(self, width, height):
        """Survey the scene's highlights and make every pixel's guide (for
        render_pass(flags=PASS_HIGHLIGHT) at this size)."""
        st = HighlightStats()
        _check(library().kdn_highlight_guides(self._h, width, height, C.byref(st)), self._h)
        return st

    def highlight_guide_at(self, x, y):
        """Pixel (x, y)'s packed guide (HG_FLOATS float32)."""
        g = np.zeros(HG_FLOATS, np.float32)
        _check(library().kdn_highlight_guide_at(self._h, x, y, _ptr(g)), self._h)
        return g

    def highlight_density(self, packed, place, normal, out, dirs):
        """The lumps' density of directions [n, 3] at a vertex."""
        g = np.ascontiguousarray(packed, np.float32)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        v = [np.ascontiguousarray(a, np.float32) for a in (place, normal, out)]
        r = np.zeros(len(d), np.float32)
        _check(library().kdn_highlight_density(self._h, _ptr(g), _ptr(v[0]), _ptr(v[1]), _ptr(v[2]), _ptr(d), len(d), _ptr(r)),
               self._h)
        return r

    def highlight_draw(self, packed, place, numbers):
        """Directions [n, 3] drawn from the lumps by (pick, u, v) [n, 3]."""
        g = np.ascontiguousarray(packed, np.float32)
        q = np.ascontiguousarray(numbers, np.float32).reshape(-1, 3)
        x = np.ascontiguousarray(place, np.float32)
        r = np.zeros((len(q), 3), np.float32)
        _check(library().kdn_highlight_draw(self._h, _ptr(g), _ptr(x), _ptr(q), len(q), _ptr(r)), self._h)
        return r

    def highlight_chains(self, queries):
        """Chains (CHAIN_DTYPE) walked on the device for queries
        (CHAIN_QUERY_DTYPE) over the uploaded scene's leaves."""
        q = np.ascontiguousarray(queries, CHAIN_QUERY_DTYPE)
        r = np.zeros(len(q), CHAIN_DTYPE)
        _check(library().kdn_highlight_chains(self._h, _ptr(q), len(q), _ptr(r)), self._h)
        return r

    # ---- the learned guide ----
    def guide_reset(self, seed=1, lower=None, size=None):
        lo = (C.c_float * 3)(*[float(x) for x in lower]) if lower is not None else None
        sz = (C.c_float * 3)(*[float(x) for x in size]) if size is not None else None
        _check(library().kdn_guide_reset(self._h, seed, lo, sz), self._h)

    def guide_configure(self, learning_rate=1e-2, share=0.3, epochs=2.0, min_steps=16, max_steps=64, max_resolution=128,
                        train_paths=1 << 20, learn_select=False, product=True):
        cfg = GuideConfig(learning_rate, share, epochs, min_steps, max_steps, max_resolution, train_paths,
                          1 if learn_select else 0, 1 if product else 0, 0)
        _check(library().kdn_guide_configure(self._h, C.byref(cfg)), self._h)

    def guide_info(self):
        info = GuideInfo()
        _check(library().kdn_guide_describe(self._h, C.byref(info)), self._h)
        return info

    def guide_params(self):
        """(grid [16, 2^19, 2], mlp [4, 64, 64]) FP32."""
        info = self.guide_info()
        grid = np.zeros(info.grid_params, np.float32)
        mlp = np.zeros(info.mlp_params, np.float32)
        _check(library().kdn_guide_params(self._h, _ptr(grid), _ptr(mlp)), self._h)
        return grid.reshape(16, -1, 2), mlp.reshape(4, 64, 64)

    def guide_set_params(self, grid=None, mlp=None):
        g = np.ascontiguousarray(grid, np.float32) if grid is not None else None
        m = np.ascontiguousarray(mlp, np.float32) if mlp is not None else None
        _check(library().kdn_guide_set_params(self._h, _ptr(g) if g is not None else None,
                                              _ptr(m) if m is not None else None), self._h)

    def guide_infer(self, queries):
        """queries [n, 10] (point, out, normal, nm) -> (lobes [n, 8, 5]
        (axis, sharpness, share), selection probabilities [n])."""
        q = np.ascontiguousarray(queries, np.float32)
        out = np.zeros((len(q), 8, 5), np.float32)
        sel = np.zeros(len(q), np.float32)
        _check(library().kdn_guide_infer(self._h, _ptr(q), len(q), _ptr(out), _ptr(sel)), self._h)
        return out, sel

    def guide_gradients(self, samples, grid=True):
        """samples [n, 16] -> (loss, grid grad [16, 2^19, 2] or None, mlp grad [4, 64, 64])."""
        s = np.ascontiguousarray(samples, np.float32)
        mg = np.zeros((4, 64, 64), np.float32)
        gg = np.zeros((16, 1 << 19, 2), np.float32) if grid else None
        loss = C.c_double()
        _check(library().kdn_guide_gradients(self._h, _ptr(s), len(s), _ptr(gg) if grid else None, _ptr(mg),
                                             C.byref(loss)), self._h)
        return loss.value, gg, mg

    def guide_train(self, samples, steps, batch):
        s = np.ascontiguousarray(samples, np.float32)
        losses = np.zeros(steps, np.float64)
        _check(library().kdn_guide_train(self._h, _ptr(s), len(s), steps, batch, _ptr(losses)), self._h)
        return losses

    def image(self, width, height):
        rgb = np.zeros((height, width, 3), np.float32)
        _check(library().kdn_read_image(self._h, _ptr(rgb)), self._h)
        return rgb


def encode_image(rgb, exposure=1.0, curve=CURVE_FILMIC):
    """A linear RGB image [height, width, 3] graded to bytes as the PPM holds them."""
    img = np.ascontiguousarray(rgb, np.float32)
    h, w, _ = img.shape
    out = np.zeros((h, w, 3), np.uint8)
    _check(library().kdn_image_encode(_ptr(img), w, h, exposure, curve, _ptr(out)))
    return out


def write_ppm(rgb, path, exposure=1.0, curve=CURVE_FILMIC):
    img = np.ascontiguousarray(rgb, np.float32)
    h, w, _ = img.shape
    _check(library().kdn_image_write_ppm(_ptr(img), w, h, exposure, curve, os.fsencode(path)))


def write_pfm(rgb, path):
    img = np.ascontiguousarray(rgb, np.float32)
    h, w, _ = img.shape
    _check(library().kdn_image_write_pfm(_ptr(img), w, h, os.fsencode(path)))


class Server:
    """Many render jobs on one device (kdn_serve_*): each step() is one
    integrator round over the admitted jobs' next slices."""

    def __init__(self, device, budget_bytes=64 << 30, max_paths=1 << 23):
        self._dev = device
        self._h = C.c_void_p()
        self._scenes = {}  # keeps the scenes of submitted jobs alive
        self._sizes = {}
        rc = library().kdn_serve_open(device.handle, budget_bytes, max_paths, C.byref(self._h))
        if rc != 0:
            raise KodanukiError(f"kodanuki error {rc}: {library().kdn_serve_error(None).decode()}")

    def _check(self, rc):
        if rc != 0:
            raise KodanukiError(f"kodanuki error {rc}: {library().kdn_serve_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            library().kdn_serve_close(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, scene, width, height, spp, seed=1, surfaces=7, min_depth=4, depth=8, guided=False, learned=False):
        """A job; guided: its first vertices sampled with the plain guide
        (a survey of its image when it is admitted); learned: its vertices
        sampled with a learned guide of its own, trained on its own paths
        after each round (one slice of the job a round)."""
        flags = (JOB_PLAIN_GUIDE if guided else 0) | (JOB_LEARNED_GUIDE if learned else 0)
        d = JobDesc(scene.handle, width, height, spp, seed, surfaces, min_depth, depth, flags, 0)
        job = C.c_uint64()
        self._check(library().kdn_serve_submit(self._h, C.byref(d), C.byref(job)))
        self._scenes[job.value] = scene
        self._sizes[job.value] = (width, height)
        return job.value

    def step(self):
        st = ServeStats()
        self._check(library().kdn_serve_step(self._h, C.byref(st)))
        return st

    def status(self, job):
        st = JobStatus()
        self._check(library().kdn_serve_status(self._h, job, C.byref(st)))
        return st

    def job_error(self, job):
        return library().kdn_serve_job_error(self._h, job).decode()

    def result(self, job):
        w, h = self._sizes[job]
        rgb = np.zeros((h, w, 3), np.float64)
        sq = np.zeros((h, w, 3), np.float64)
        n = np.zeros((h, w), np.uint32)
        self._check(library().kdn_serve_result(self._h, job, _ptr(rgb), _ptr(sq), _ptr(n)))
        return rgb, sq, n

    def release(self, job):
        self._check(library().kdn_serve_release(self._h, job))
        self._scenes.pop(job, None)
        self._sizes.pop(job, None)

    def run(self, until=None):
        """Step until every submitted job (or the given ones) has ended."""
        jobs = list(self._sizes) if until is None else list(until)
        steps = 0
        while any(self.status(j).state in (JOB_WAITING, JOB_RUNNING) for j in jobs):
            self.step()
            steps += 1
        return steps


def make_rays(origins, directions, far=np.inf, inside=0, within=NO_LEAF):
    r = np.zeros(len(origins), RAY_DTYPE)
    r["origin"] = origins
    r["direction"] = directions
    r["far_"] = far
    r["inside"] = inside
    r["within"] = within
    return r
