"""Seeded synthetic scenes, written as .sdf text for the scene loader.

Random numbers are SplitMix64 (the reference's mixedBits,
source/math/random.cpp:8-13) over a counter, so a scene is the same for a
seed on every machine. The shapes are the standard.sdf primitives
(sphere, box, torus, plane); the statements are those the loader reads.

Workloads:
- plain(32):   32 primitives, smooth-merged (k = 0.05) over a ground plane,
               matte albedos U[0.2, 0.8], one quad light, 256x256.
- production(512): 512 primitives in groups nested by cut/clip/merge to
               depth 4, 40% matte / 30% rough metal / 30% dispersive glass,
               four quad lights and a gradient sky, 1920x1080.
- glass_heavy: the production scene with 60% glass (config 2).
- serving_pool: eight of the above, 32 to 4096 primitives.
- window_room: a room lit by the sky through a small ceiling window.
"""
import math

MASK = (1 << 64) - 1


def mixed_bits(v):
    v = (v + 0x9E3779B97F4A7C15) & MASK
    v = ((v ^ (v >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    v = ((v ^ (v >> 27)) * 0x94D049BB133111EB) & MASK
    return v ^ (v >> 31)


class Rng:
    def __init__(self, seed):
        self.stream = mixed_bits(seed)
        self.n = 0

    def unit(self):
        self.n += 1
        return (mixed_bits((self.stream + self.n) & MASK) >> 11) * 2.0 ** -53

    def uniform(self, lo, hi):
        return lo + (hi - lo) * self.unit()

    def pick(self, weights):
        u = self.unit() * sum(weights)
        for i, w in enumerate(weights):
            if u < w:
                return i
            u -= w
        return len(weights) - 1


def _v(x):
    return "[" + ", ".join(f"{c:.6g}" for c in x) + "]"


def _primitive(rng, scale=1.0):
    kind = rng.pick([1, 1, 1])
    if kind == 0:
        r = rng.uniform(0.1, 0.5) * scale
        return f"sphere({r:.6g})", r
    if kind == 1:
        size = [rng.uniform(0.2, 0.8) * scale for _ in range(3)]
        r = min(0.05 * scale, min(size) / 4)
        return f"box({_v(size)}, {r:.6g})", max(size) / 2
    big = rng.uniform(0.15, 0.4) * scale
    small = rng.uniform(0.04, 0.12) * scale
    return f"torus({big:.6g}, {small:.6g})", big + small


def _material(rng, name, kinds, glass=True):
    k = rng.pick(kinds)
    if k == 2 and not glass:
        # the same draws as a glass (so the rest of the scene is the same),
        # painted as a white matte
        rng.uniform(1.4, 1.7)
        rng.uniform(0.003, 0.01)
        return f"let {name} = matte({_v([0.9, 0.9, 0.9])});"
    if k == 0:
        a = [rng.uniform(0.2, 0.8) for _ in range(3)]
        return f"let {name} = matte({_v(a)});"
    if k == 1:
        a = [rng.uniform(0.5, 0.95) for _ in range(3)]
        alpha = rng.uniform(0.05, 0.5)
        return f"let {name} = metal({_v(a)}, roughness = {math.sqrt(alpha):.6g});"
    A = rng.uniform(1.4, 1.7)
    B = rng.uniform(0.003, 0.01)
    return f"let {name} = glass(refraction = {A:.6g}, dispersion = {B:.6g});"


def plain(seed=0, count=32, glass=0.0, width=256, height=256, samples=16, surfaces=7, spread=False):
    """The plain scene: `count` primitives smooth-merged over a plane.
    They sit on a square grid 3.6 wide; with spread=True the grid keeps the
    spacing of 32 primitives (0.6) and grows with the count, and the
    camera, light and region grow with it (the same density of shapes, so
    the same length of per-cell programs, at any count)."""
    rng = Rng(seed)
    side = max(1, int(math.ceil(math.sqrt(count))))
    spacing = 0.6 if spread else 3.6 / side
    f = max(1.0, spacing * side / 3.6)
    out = ['import "standard.sdf";',
           f"render {{ width = {width}; height = {height}; seed = {seed + 1}; samples = {samples}; surfaces = {surfaces}; "
           f"far = {4 * f:.6g}; }}",
           f"camera {{ eye = {_v([0, -6.5 * f, 3.2 * f])}; target = [0, 0, 0.45]; up = [0, 0, 1]; fov = 0.75; }}",
           f"light quad {{ at = {_v([-0.75 * f, -1.25 * f, 4 * f])}; along = {_v([0, 1.5 * f, 0])}; "
           f"across = {_v([1.5 * f, 0, 0])}; power = {_v([40 * f * f, 38 * f * f, 34 * f * f])}; }}"]
    kinds = [1.0 - glass, 0.0, glass] if glass > 0 else [1.0, 0.0, 0.0]
    for i in range(count):
        out.append(_material(rng, f"m{i}", kinds))
    out.append("let floor = matte([0.55, 0.55, 0.5]);")
    out.append("merge(0.05) {")
    half = spacing * side / 2
    for i in range(count):
        shape, reach = _primitive(rng)
        gx, gy = i % side, i // side
        x = -half + spacing * (gx + 0.5) + rng.uniform(-0.15, 0.15)
        y = -half + spacing * (gy + 0.5) + rng.uniform(-0.15, 0.15)
        z = reach * rng.uniform(0.6, 1.0)
        ang = rng.uniform(0, math.pi)
        out.append(f"\tpaint(m{i}) at({_v([x, y, z])}) turn(z, {ang:.6g}) {shape};")
    out.append("}")
    out.append("paint(floor) plane([0, 0, 1], 0);")
    return "\n".join(out) + "\n"


def glass_heavy(seed=0, **kw):
    """Config 2's scene: the production scene (config 1) with glass-heavy
    caustics, 60% of its materials dielectric with Cauchy dispersion, the
    rest matte and rough metal alike."""
    return production(seed, kinds=(0.2, 0.2, 0.6), **kw)


def _nest(rng, depth, leaves, materials):
    """A CSG nest of `leaves` primitives to `depth` levels of joins."""
    if depth == 0 or leaves == 1:
        shape, _ = _primitive(rng, 0.8)
        m = materials[int(rng.unit() * len(materials))]
        off = [rng.uniform(-0.25, 0.25), rng.uniform(-0.25, 0.25), rng.uniform(-0.1, 0.2)]
        return [f"paint({m}) at({_v(off)}) {shape};"]
    join = ["merge", "cut", "clip"][rng.pick([0.5, 0.3, 0.2])]
    radius = [0, 0.02, 0.05][rng.pick([0.4, 0.4, 0.2])]
    head = f"{join}({radius:.6g}) {{" if radius > 0 else f"{join} {{"
    left = leaves // 2
    body = _nest(rng, depth - 1, left, materials) + _nest(rng, depth - 1, leaves - left, materials)
    return [head] + ["\t" + line for line in body] + ["}"]


def production(seed=0, count=512, width=1920, height=1080, samples=64, surfaces=13, glass=True, kinds=(0.4, 0.3, 0.3)):
    """The production scene: groups of 4 primitives nested to depth 4, its
    materials matte, rough metal and dispersive glass in the shares `kinds`.
    With glass=False its glass materials are white mattes (same geometry)."""
    rng = Rng(seed + 1000)
    out = ['import "standard.sdf";',
           f"render {{ width = {width}; height = {height}; seed = {seed + 1}; samples = {samples}; surfaces = {surfaces}; far = 8; }}",
           "camera { eye = [0, -13, 6.5]; target = [0, 0, 0.3]; up = [0, 0, 1]; fov = 0.8; }",
           "sky gradient { zenith = [0.25, 0.35, 0.6]; horizon = [0.6, 0.6, 0.62]; ground = [0.08, 0.08, 0.08]; }"]
    for k, (x, y) in enumerate([(-4, -4), (3, -4), (-4, 3), (3, 3)]):
        p = 30 + 10 * k
        out.append(f"light quad {{ at = {_v([x, y, 5])}; along = [0, 1.2, 0]; across = [1.2, 0, 0]; power = {_v([p, p * 0.95, p * 0.9])}; }}")
    nm = 48
    for i in range(nm):
        out.append(_material(rng, f"m{i}", list(kinds), glass))
    out.append("let floor = matte([0.5, 0.5, 0.48]);")
    groups = count // 4
    side = int(math.ceil(math.sqrt(groups)))
    spacing = 12.0 / side
    mats = [f"m{i}" for i in range(nm)]
    for g in range(groups):
        gx, gy = g % side, g // side
        x = -6 + spacing * (gx + 0.5)
        y = -6 + spacing * (gy + 0.5)
        body = _nest(rng, 4, 4, mats)
        out.append(f"at({_v([x, y, 0.35])}) turn(z, {rng.uniform(0, math.pi):.6g}) " + body[0])
        out.extend(body[1:])
    out.append("paint(floor) plane([0, 0, 1], 0);")
    return "\n".join(out) + "\n"


def stress(seed=0, count=4096, grid=64, width=3840, height=2160, samples=256, surfaces=32, glass=True):
    """The stress scene: the production scene grown to `count` primitives,
    with a canopy of `grid` x `grid` domain-repeated instances above it
    (the reference's repeat(count, step), syntax/evaluator.cpp:2457-2577):
    each a thin-walled glass bubble (a shell 0.004 thick, cut from a sphere)
    holding a nested dielectric core of another refraction. With
    glass=False the shells and cores (and the production scene's glass) are
    white and grey mattes, the geometry the same."""
    base = production(seed, count=count, width=width, height=height, samples=samples, surfaces=surfaces, glass=glass)
    lines = base.rstrip("\n").split("\n")
    rng = Rng(seed + 2000)
    a_shell, b_shell = rng.uniform(1.4, 1.7), rng.uniform(0.003, 0.01)
    a_core, b_core = rng.uniform(1.3, 1.6), rng.uniform(0.003, 0.01)
    if glass:
        mats = [f"let shell = glass(refraction = {a_shell:.6g}, dispersion = {b_shell:.6g});",
                f"let core = glass(refraction = {a_core:.6g}, dispersion = {b_core:.6g});"]
    else:
        mats = ["let shell = matte([0.9, 0.9, 0.9]);", "let core = matte([0.6, 0.6, 0.62]);"]
    step = 11.52 / grid
    r = 0.4 * step
    canopy = [f"at({_v([-5.76 + step / 2, -5.76 + step / 2, 3.2])}) repeat(count = [{grid}, {grid}, 1], step = {_v([step, step, 0])}) {{",
              f"\tpaint(shell) cut {{ sphere({r:.6g}); sphere({r - 0.004:.6g}); }}",
              f"\tpaint(core) sphere({0.6 * r:.6g});",
              "}"]
    # materials before the shapes, the canopy before the floor
    return "\n".join(lines[:-1] + mats + canopy + lines[-1:]) + "\n"


def window_room(width=64, height=48, samples=16, surfaces=6):
    """A closed grey room lit only by a bright sky through a small square
    window in its ceiling (no lamps): every pixel's light comes through
    0.25 m^2 of opening that BSDF sampling finds rarely, the case path
    guiding is for."""
    return f"""import "standard.sdf";
render {{ width = {width}; height = {height}; seed = 1; samples = {samples}; surfaces = {surfaces}; far = 4; }}
camera {{ eye = [-1.6, -1.6, 1.1]; target = [0.6, 0.6, 0.7]; up = [0, 0, 1]; fov = 1.1; }}
sky gradient {{ zenith = [12, 12, 12]; horizon = [6, 6, 6]; ground = [0, 0, 0]; }}
let wall = matte([0.75, 0.75, 0.72]);
let red = matte([0.75, 0.2, 0.18]);
paint(wall) cut {{ at([0, 0, 1]) box([4.4, 4.4, 2.2]); at([0, 0, 1]) box([4, 4, 2]); at([0.8, 0.8, 2.05]) box([0.5, 0.5, 0.6]); }}
paint(red) at([0.3, -0.2, 0.35]) sphere(0.35);
"""


_HIGHLIGHT_LAMPS = {
    "quad": "light quad { at = [-1.35, -0.15, 1]; along = [0, 0.3, 0]; across = [0.3, 0, 0]; power = [40, 40, 38]; }",
    # the same power from a ball where the quad's centre is
    "sphere": "light sphere { at = [-1.2, 0, 1]; radius = 0.12; power = [40, 40, 38]; }",
}


def highlight_room(width=64, height=48, samples=16, surfaces=6, roughness=0.15, lamp="quad"):
    """A closed grey room lit by one small lamp that faces the floor, where
    a polished steel plate mirrors it onto the far wall: the wall above the
    lamp's height sees no lamp directly, and its light comes off the plate
    (a highlight, as the plain guide surveys them) or off the floor."""
    return f"""import "standard.sdf";
render {{ width = {width}; height = {height}; seed = 1; samples = {samples}; surfaces = {surfaces}; far = 4; }}
camera {{ eye = [-1.7, -1.7, 1.4]; target = [2, 0.3, 1.7]; up = [0, 0, 1]; fov = 1.1; }}
{_HIGHLIGHT_LAMPS[lamp]}
let wall = matte([0.75, 0.75, 0.72]);
let steel = metal([0.95, 0.93, 0.9], roughness = {roughness:.6g});
paint(wall) cut {{ at([0, 0, 1.5]) box([4.4, 4.4, 3.4]); at([0, 0, 1.5]) box([4, 4, 3]); }}
paint(steel) at([0, 0, 0.01]) box([1.2, 1.2, 0.02]);
"""


def stopping_table(width=64, height=40, tolerance=6, samples=64, batch=8):
    """A matte table top with a plastic and a metal ball under a quad lamp
    and a dim sky: quiet floor and sky pixels, noisy highlights and
    shadow edges, the spread adaptive stopping divides its samples by."""
    return f"""import "standard.sdf";
render {{ width = {width}; height = {height}; seed = 1; tolerance = {tolerance}; samples = {samples}; batch = {batch}; surfaces = 6; far = 8; }}
camera {{ eye = [0, -4.2, 1.6]; target = [0, 0, 0.35]; up = [0, 0, 1]; fov = 0.7; }}
sky gradient {{ zenith = [0.25, 0.3, 0.45]; horizon = [0.45, 0.45, 0.45]; ground = [0.1, 0.1, 0.1]; }}
light quad {{ at = [0.8, -1.4, 2.2]; along = [0.4, 0, 0]; across = [0, 0.4, 0]; power = [9, 9, 8]; }}
paint(matte([0.6, 0.6, 0.58])) at([0, 0, -0.1]) box([6, 6, 0.2]);
paint(plastic([0.8, 0.25, 0.1], roughness = 0.3)) at([-0.6, 0, 0.5]) sphere(0.5);
paint(metal([0.9, 0.8, 0.5], roughness = 0.1)) at([0.65, 0.3, 0.4]) sphere(0.4);
"""


SERVING_COUNTS = (32, 64, 128, 256, 512, 1024, 2048, 4096)


def serving_pool(glass=True):
    """Eight seeded scenes of 32 to 4096 primitives for the serving stream:
    (text, octree depth) pairs, the plain scene spread at the density of 32
    primitives (octree depth 6, 7, 8) up to 128 primitives, the production
    scene (depth 8) from 256 on."""
    pool = []
    for i, n in enumerate(SERVING_COUNTS):
        if n <= 128:
            # smooth merges of more shapes keep more of them per cell: a
            # level deeper per doubling keeps every surface cell's program
            # within the march's 384 registers
            pool.append((plain(100 + i, count=n, spread=True), {32: 6, 64: 7, 128: 8}[n]))
        else:
            pool.append((production(100 + i, count=n, glass=glass), 8))
    return pool


def tiny(seed=0, glass=True):
    """A handful of shapes for quick checks: a sphere, a box and a torus.
    With glass=False the torus is a white matte instead of glass."""
    text = _TINY if glass else _TINY.replace("let pane = glass(refraction = 1.5, dispersion = 0.005);",
                                             "let pane = matte([0.9, 0.9, 0.9]);")
    return text


_TINY = """import "standard.sdf";
render { width = 64; height = 48; seed = 1; samples = 4; surfaces = 4; far = 3; }
camera { eye = [0, -5, 2]; target = [0, 0, 0.4]; up = [0, 0, 1]; fov = 0.8; }
light quad { at = [-0.5, -0.5, 3]; along = [0, 1, 0]; across = [1, 0, 0]; power = [20, 20, 20]; }
let red = matte([0.7, 0.2, 0.2]);
let steel = metal([0.9, 0.9, 0.9], roughness = 0.4);
let pane = glass(refraction = 1.5, dispersion = 0.005);
let grey = matte([0.5, 0.5, 0.5]);
merge(0.1) {
	paint(red) at([-0.7, 0, 0.4]) sphere(0.4);
	paint(steel) at([0.1, 0.2, 0.3]) turn(z, 0.5) box([0.5, 0.6, 0.5], 0.05);
}
paint(pane) at([0.9, -0.3, 0.25]) torus(0.3, 0.1);
cut(0.02) { paint(grey) at([0, 0.9, 0.3]) box([0.6, 0.6, 0.6], 0.02); at([0.2, 0.8, 0.5]) sphere(0.3); }
paint(grey) plane([0, 0, 1], 0);
"""
