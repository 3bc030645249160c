"""The render entry point on the GPU: `workflow.py render scene.sdf config.toml`
(target/workflow.py:181-185) for the MI355X path.

    python -m kodanuki_amd.render render scene.sdf config.toml [--out DIR] [--depth N] [--device N]

The config is the reference's TOML render config (target/configs/*.toml):
[settings], [viewer] and [plugins], checked as workflow.py checks them
(config_of, plugins_of: known sections and keys only, true/false plugins,
exactly one geometry, execution, backend and stopping), with `hip` added to
the geometry, execution and backend members (the keys a maintainer adds to
pluginKeys, core/options.cpp:36-47, and to KEYS, workflow.py:21-30). The GPU
path runs with backend/hip; a config that chooses backend/cpu is the
reference's own renderer and is refused here.

What each setting and plugin does on the GPU path:
- width, height, samples, tolerance, batch, split, surfaces, seed: as the
  renderer's --size/--samples/... flags (workflow.py:103-113), over the
  scene file's render block; split must be 1 (every sample draws its own
  first vertex); threads is the CPU backend's and is ignored.
- stopping/confidence_interval: the frame is drawn as the reference's
  frame draws it (core/frame.cpp:28-62): a trial of two batches in every
  pixel, each pixel judged and allotted on the device, then the rest;
  stopping/fixed_count: one pass of `samples`.
- sampler/sobol: an eye path's first 28 coordinates off the Owen-scrambled
  Sobol sequence; sampler/highlight_guide: the plain guide's survey and
  pixel guides steer each path's first vertex.
- reconstruction/atrous: the shown image is the a-trous filter of the
  frame's history (its G-buffer from primary rays); reconstruction/mean (or
  none): the pixel means.
- sink/file: the shown image as output/<scene>.ppm (graded with the scene's
  exposure and curve) and output/<scene>.pfm (linear), as the renderer's
  --out and --pfm write them.
- geometry/hip: the interval-cut octree of the scene (depth --depth, 8 by
  default); execution/hip: rays marched on the device. The reference's
  tree-build, detail, strategy and feature plugins shape its CPU tree and
  CPU transport; the GPU path takes their keys and draws the same
  estimator without them.
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

NUMBERS = ("tolerance", "samples", "batch", "split", "surfaces", "collisions", "seed", "threads")
SETTINGS = ("width", "height", "faces", "smooth_glass") + NUMBERS
VIEWER = ("internal_width", "internal_height")
KEYS = {
    "geometry": ("tree", "sphere_trace", "hip"),
    "execution": ("single", "packets", "hip"),
    "backend": ("cpu", "hip"),
    "stopping": ("fixed_count", "confidence_interval"),
    "detail": ("focus_on_view", "pilot", "split_hidden_primitives", "split_crowded_leaves"),
    "build": ("split_beside_finer", "cut_long_programs", "face_slabs", "shared_coefficients", "shared_subtrees"),
    "strategy": ("join", "connect"),
    "sampler": ("sobol", "highlight_guide", "each_lamp_first", "sky_map", "glowing_surfaces", "toward_mirrors"),
}
LOOP_KEYS = {
    "source": ("scene", "flight"),
    "sink": ("file", "frames", "window"),
    "observer": ("progress", "metrics"),
    "reconstruction": ("mean", "temporal", "atrous", "upsample", "upscale"),
}
ONE = ("geometry", "execution", "backend", "stopping")
BUILT = ("build/faced_hits", "feature/clouds", "feature/path_rules", "feature/coats")
ALWAYS = ("feature/absorption", "feature/atmosphere", "feature/dispersion", "feature/scattering", "feature/sky_layers",
          "transport/paths")
SWITCHED = {f"{key}/{member}": key for key, members in {**KEYS, **LOOP_KEYS}.items() for member in members}
PLUGINS = sorted(list(SWITCHED) + list(BUILT) + list(ALWAYS))
# the reference's defaults (workflow.py:41-44), the hip members off
OFF = ("geometry/sphere_trace", "execution/single", "stopping/fixed_count", "strategy/connect", "source/flight", "sink/frames",
       "sink/window", "observer/metrics", "reconstruction/temporal", "reconstruction/atrous", "reconstruction/upsample",
       "reconstruction/upscale", "geometry/hip", "execution/hip", "backend/hip")
DEFAULTS = {name: name not in OFF for name in PLUGINS}


class ConfigError(ValueError):
    """A config the render entry point refuses, with the reason
    (workflow.py's refused() messages)."""


def _checked(path, block, known):
    for key in block:
        if key not in known:
            raise ConfigError(f"{path}: unknown key {key}, known: {', '.join(known)}")
    return block


def plugins_of(path, block):
    """plugins_of (workflow.py:62-72) over the members with hip."""
    plugins = {**DEFAULTS, **_checked(path, block, PLUGINS)}
    for name, on in plugins.items():
        if not isinstance(on, bool):
            raise ConfigError(f'{path}: plugins."{name}" takes true or false')
        if not on and name in ALWAYS:
            raise ConfigError(f'{path}: plugins."{name}" cannot be left out')
    for key in ONE:
        if [plugins[name] for name, of in SWITCHED.items() if of == key].count(True) != 1:
            raise ConfigError(f"{path}: exactly one {key} plugin is on")
    return plugins


def _toml(path):
    try:
        import tomllib
    except ImportError:  # Python < 3.11: the same parser as a package
        import tomli as tomllib
    with open(path, "rb") as f:
        return tomllib.load(f)


def config_of(path):
    """config_of (workflow.py:75-83): the config's sections, checked."""
    if not os.path.isfile(path):
        raise ConfigError(f"{path}: no such config")
    config = _checked(path, _toml(path), ("settings", "viewer", "plugins"))
    return {"settings": _checked(path, config.get("settings", {}), SETTINGS),
            "viewer": _checked(path, config.get("viewer", {}), VIEWER),
            "plugins": plugins_of(path, config.get("plugins", {}))}


def gpu_plan(config):
    """What the GPU path draws for a checked config: a dict of the choices
    the render below follows. Refuses configs the GPU path cannot draw."""
    settings, plugins = config["settings"], config["plugins"]
    if not plugins["backend/hip"]:
        raise ConfigError("the GPU path runs with backend/hip (backend/cpu is the reference's own renderer)")
    if not plugins["geometry/hip"] or not plugins["execution/hip"]:
        raise ConfigError("backend/hip draws with geometry/hip and execution/hip")
    if settings.get("split", 1) != 1:
        raise ConfigError("split: every sample draws its own first vertex on the GPU path (split 1)")
    if settings.get("smooth_glass", "exact") != "exact":
        raise ConfigError("smooth_glass: the GPU path's smooth glass is exact")
    for name in ("reconstruction/temporal", "reconstruction/upsample", "reconstruction/upscale", "sink/frames", "sink/window",
                 "source/flight"):
        if plugins[name]:
            raise ConfigError(f'plugins."{name}" belongs to the viewer or a sequence; a render draws one still frame')
    return {"stopping": "confidence_interval" if plugins["stopping/confidence_interval"] else "fixed_count",
            "sobol": plugins["sampler/sobol"], "highlight_guide": plugins["sampler/highlight_guide"],
            "denoise": plugins["reconstruction/atrous"], "file": plugins["sink/file"],
            "settings": {k: v for k, v in settings.items() if k in ("width", "height") + NUMBERS}}


def render(scene_path, config_path, out_dir=None, depth=8, device=0, dev=None):
    """Draw the scene's still frame as the config says; returns (shown rgb
    [h, w, 3] float32, paths written, the plan). dev: an open Device to draw
    on (one is opened on `device` otherwise)."""
    import numpy as np

    import kodanuki_amd as k
    plan = gpu_plan(config_of(config_path))
    scene = k.Scene.from_file(scene_path)
    cam, rs = scene.settings()
    s = plan["settings"]
    width, height = int(s.get("width", rs.width)), int(s.get("height", rs.height))
    seed = int(s.get("seed", rs.seed))
    surfaces = int(s.get("surfaces", rs.surfaces))
    for key in ("samples", "batch"):
        if key in s:
            setattr(rs, key, int(s[key]))
    if "tolerance" in s:
        rs.tolerance = float(s["tolerance"])
    flags = (k.PASS_SOBOL if plan["sobol"] else 0)
    lower, size = scene.region(depth)
    own = dev is None
    dev = dev or k.Device(device)
    try:
        dev.upload(scene)
        dev.build_octree(lower, size, depth)
        if plan["highlight_guide"]:
            dev.highlight_guides(width, height)
            flags |= k.PASS_HIGHLIGHT
        if plan["stopping"] == "confidence_interval":
            stopping = k.stopping_of(rs)
            dev.render_stopped(width, height, stopping, seed=seed, surfaces=surfaces, flags=flags)
        else:
            stopping = k.stopping_of(rs)
            dev.render_pass(width, height, stopping.samples, seed=seed, surfaces=surfaces, flags=flags | k.PASS_ESTIMATES,
                            batch=stopping.batch)
        if plan["denoise"]:
            shown = dev.denoise_accumulation(width, height)
        else:
            est = dev.estimates(width, height)
            taken = np.maximum(est["taken"], 1).astype(np.float64)[..., None]
            shown = (est["total"] / taken).astype(np.float32)
    finally:
        if own:
            dev.close()
    written = []
    if plan["file"]:
        out_dir = out_dir or os.path.join(os.getcwd(), "output")
        os.makedirs(out_dir, exist_ok=True)
        stem = os.path.splitext(os.path.basename(scene_path))[0]
        ppm, pfm = os.path.join(out_dir, stem + ".ppm"), os.path.join(out_dir, stem + ".pfm")
        k.write_ppm(shown, ppm, exposure=rs.exposure, curve=rs.curve)
        k.write_pfm(shown, pfm)
        written = [ppm, pfm]
    return shown, written, plan


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m kodanuki_amd.render",
                                 description="workflow.py render scene.sdf config.toml, on the GPU path")
    ap.add_argument("command", choices=["render"])
    ap.add_argument("scene")
    ap.add_argument("config")
    ap.add_argument("--out", default=None, help="folder of the frame's .ppm and .pfm (default ./output)")
    ap.add_argument("--depth", type=int, default=8, help="the octree's depth cap (0..11)")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)
    try:
        _, written, _ = render(args.scene, args.config, args.out, args.depth, args.device)
    except ConfigError as e:
        print(e, file=sys.stderr)
        return 1
    for path in written:
        print(path)
    return 0


if __name__ == "__main__":
    sys.exit(main())
