"""Seeded ray sets for the geometry checks.

Half camera rays through jittered pixel positions, half random rays from
points outside every solid toward random points of the region, as the
reference's walk checks draw camera rays and random leaf-to-leaf rays
(target/tests/walk.cpp:181-248). A share of rays start inside a solid.
"""
import numpy as np

import kodanuki_amd as k


def camera_rays(scene, count, seed, width=None, height=None):
    cam, rs = scene.settings()
    width = width or rs.width
    height = height or rs.height
    rng = np.random.default_rng(seed)
    eye = np.array(cam.eye[:], np.float64)
    target = np.array(cam.target[:], np.float64)
    up = np.array(cam.up[:], np.float64)
    forward = (target - eye) / np.linalg.norm(target - eye)
    right = np.cross(forward, up)
    right /= np.linalg.norm(right)
    upv = np.cross(right, forward)
    reach = np.tan(cam.fov / 2)
    aspect = width / height
    x = rng.random(count) * width
    y = rng.random(count) * height
    u = 2 * x / width - 1
    v = 1 - 2 * y / height
    d = forward[None] + right[None] * (u * reach * aspect)[:, None] + upv[None] * (v * reach)[:, None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return k.make_rays(np.repeat(eye[None].astype(np.float32), count, 0), d.astype(np.float32))


def random_rays(arrays, lower, size, count, seed, inside_share=0.0):
    """Rays between random points of the region; a share start inside."""
    rng = np.random.default_rng(seed)
    lower = np.asarray(lower, np.float64)
    size = np.asarray(size, np.float64)
    origins, dirs, inside = [], [], []
    want_in = int(count * inside_share)
    tries = 0
    while len(origins) < count and tries < count * 50:
        tries += 1
        p = lower + rng.random(3) * size
        q = lower + rng.random(3) * size
        f = arrays.csg_distance(p)
        is_in = f < -0.01
        if is_in and sum(inside) >= want_in:
            continue
        if not is_in and f < 0.02:
            continue
        if not is_in and len(origins) - sum(inside) >= count - want_in:
            continue
        d = q - p
        n = np.linalg.norm(d)
        if n < 1e-3:
            continue
        origins.append(p)
        dirs.append(d / n)
        inside.append(1 if is_in else 0)
    rays = k.make_rays(np.array(origins, np.float32), np.array(dirs, np.float32))
    rays["inside"] = np.array(inside, np.uint32)
    return rays


def leaf_rays(arrays, scene, count, seed):
    """Random leaf-to-leaf rays (target/tests/walk.cpp:174-193): from a point
    within 1.5 of one primitive's placement toward a point within 0.3 of
    another's; a ray that starts inside a solid is marked inside."""
    rng = np.random.default_rng(seed)
    centres = np.array([l.origin[:] for l in scene.leaves()], np.float64)
    a = centres[rng.integers(0, len(centres), count)] + rng.uniform(-1.5, 1.5, (count, 3))
    b = centres[rng.integers(0, len(centres), count)] + rng.uniform(-0.3, 0.3, (count, 3))
    d = b - a
    n = np.linalg.norm(d, axis=1)
    keep = n > 1e-3
    a, d, n = a[keep], d[keep], n[keep]
    rays = k.make_rays(a.astype(np.float32), (d / n[:, None]).astype(np.float32))
    rays["inside"] = np.array([1 if arrays.csg_distance(p) < 0 else 0 for p in rays["origin"].astype(np.float64)], np.uint32)
    return rays


def pixel_rays(scene, count, seed, width, height):
    """Camera rays through the centres of `count` random pixels of a
    width x height image (the reference's pixelRay set, walk.cpp:227-231)."""
    cam, _ = scene.settings()
    rng = np.random.default_rng(seed)
    eye = np.array(cam.eye[:], np.float64)
    target = np.array(cam.target[:], np.float64)
    up = np.array(cam.up[:], np.float64)
    forward = (target - eye) / np.linalg.norm(target - eye)
    right = np.cross(forward, up)
    right /= np.linalg.norm(right)
    upv = np.cross(right, forward)
    reach = np.tan(cam.fov / 2)
    x = rng.integers(0, width, count) + 0.5
    y = rng.integers(0, height, count) + 0.5
    u = 2 * x / width - 1
    v = 1 - 2 * y / height
    d = forward[None] + right[None] * (u * reach * width / height)[:, None] + upv[None] * (v * reach)[:, None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return k.make_rays(np.repeat(eye[None].astype(np.float32), count, 0), d.astype(np.float32))
