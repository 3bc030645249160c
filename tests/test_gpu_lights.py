"""The light kinds on the device, against closed forms: a matte plane of
albedo a under one light, seen from above, shows (a / pi) E where E is the
light's irradiance there (aimLamp, meetLight: light.cpp:263-420):
- a point light of power P at height h: E = P cos / d^2 (its power is what
  it sends per steradian: aimPoint's arriving P / d^2);
- a sphere of radius r and power P (radiance P / (pi r^2)), wholly above
  the horizon: E = pi L (r / d)^2 cos = P cos / d^2 (aimed at through the
  cone it fills, met by the plane's own draws, MIS between the two);
- a directional light: E = P cos;
- a sun of full angle 2 alpha straight above (radiance P / (pi sin^2
  alpha)): E = pi L sin^2 alpha = P (aimed at through its cone, met by
  draws that escape into it).
Paths stop after the plane (two surfaces) and the sky is black, so the
plane's light is the direct light alone."""
import numpy as np
import pytest

import kodanuki_amd as k
from tests.conftest import EMULATION

pytestmark = pytest.mark.gpu

W = H = 24 if EMULATION else 64
SPP = 64 if EMULATION else 1024
ALBEDO, EYE_H, FOV = 0.5, 1.0, 1.0


def note(msg):
    print(msg, flush=True)


def _scene(light):
    return f"""import "standard.sdf";
render {{ width = {W}; height = {H}; seed = 5; samples = {SPP}; surfaces = 2; far = 4; }}
camera {{ eye = [0, 0, {EYE_H}]; target = [0, 0, 0]; up = [0, 1, 0]; fov = {FOV}; }}
{light}
let floor = matte([{ALBEDO}, {ALBEDO}, {ALBEDO}]);
paint(floor) plane([0, 0, 1], 0);
"""


def _plane_points():
    """The plane point each pixel's centre sees."""
    reach = np.tan(FOV / 2)
    ys, xs = np.mgrid[0:H, 0:W]
    u = 2 * (xs + 0.5) / W - 1
    v = 1 - 2 * (ys + 0.5) / H
    # forward -z, right = forward x up = (1, 0, 0)... from the camera's frame
    forward = np.array([0, 0, -1.0])
    up = np.array([0, 1.0, 0])
    right = np.cross(forward, up)
    up = np.cross(right, forward)
    d = forward[None, None] + right[None, None] * (u * reach * W / H)[..., None] + up[None, None] * (v * reach)[..., None]
    t = EYE_H / -d[..., 2]
    return np.array([0, 0, EYE_H])[None, None] + d * t[..., None]


def _render(device, light, depth=5):
    s = k.Scene(_scene(light))
    lower, size = s.region(depth)
    device.upload(s)
    device.build_octree(lower, size, depth)
    device.render_pass(W, H, SPP, seed=5, surfaces=2)
    rgb, sq, n = device.accumulation(W, H)
    mean = rgb / n[..., None]
    se = np.sqrt(np.maximum(sq / n[..., None] - mean ** 2, 0) / n[..., None])
    return mean[..., 1], se[..., 1]


def _check(name, got, se, want):
    z = np.abs(got - want) / (se + 1e-4 * want)
    rel = abs(got.mean() / want.mean() - 1)
    note(f"[lights] {name}: mean {got.mean():.5f} closed form {want.mean():.5f} (rel {rel:.2e}), "
         f"within 4 se {np.mean(z < 4):.4f}")
    assert rel < (0.02 if EMULATION else 0.005)
    assert np.mean(z < 4) > 0.99


def test_point_light(device):
    P, h = 4.0, 2.0
    got, se = _render(device, f"light point {{ at = [0, 0, {h}]; power = [{P}, {P}, {P}]; }}")
    p = _plane_points()
    d2 = p[..., 0] ** 2 + p[..., 1] ** 2 + h * h
    _check("point", got, se, ALBEDO / np.pi * P * h / d2 ** 1.5)


def test_sphere_light(device):
    P, h, r = 4.0, 2.0, 0.3
    got, se = _render(device, f"light sphere {{ at = [0, 0, {h}]; radius = {r}; power = [{P}, {P}, {P}]; }}")
    p = _plane_points()
    d2 = p[..., 0] ** 2 + p[..., 1] ** 2 + h * h
    _check("sphere", got, se, ALBEDO / np.pi * P * h / d2 ** 1.5)


def test_directional_light(device):
    P = 2.0
    direction = np.array([0.3, 0.2, 1.0])
    got, se = _render(device, f"light directional {{ direction = [{direction[0]}, {direction[1]}, {direction[2]}]; "
                              f"power = [{P}, {P}, {P}]; }}")
    cos = direction[2] / np.linalg.norm(direction)
    _check("directional", got, se, np.full((H, W), ALBEDO / np.pi * P * cos))


def test_sun(device):
    P = 2.0
    got, se = _render(device, f"light sun {{ direction = [0, 0, 1]; angle = 0.4; power = [{P}, {P}, {P}]; }}")
    _check("sun", got, se, np.full((H, W), ALBEDO / np.pi * P))


def test_moon_is_a_sun(device):
    """A moon is aimed at and met as a sun (light.cpp treats the two alike)."""
    P = 2.0
    got, se = _render(device, f"light moon {{ direction = [0, 0, 1]; angle = 0.4; power = [{P}, {P}, {P}]; }}")
    _check("moon", got, se, np.full((H, W), ALBEDO / np.pi * P))
