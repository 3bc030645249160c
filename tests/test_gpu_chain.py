"""The highlight survey's chain walk on the device (kdn_highlight_chains,
FP64 hyper-dual evaluation of each leaf's own program) against the
reference's own walkedChain (chain.cpp:507-519, oracle/_ref) on the same
queries: a sphere, a turned rounded box and a turned torus, each toward a
point, a quad, a sphere, a sun and a directional lamp from 24 floor points.

Both walk in FP64 with the same steps, so they settle alike but for
rounding: the same chains settle, at points within 1e-7, with spreads and
motions within 1e-5 of their size."""
import numpy as np
import pytest

import kodanuki_amd as k
from tests import chain_cases as CC

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)


def test_device_chains_match_the_reference(device):
    lib = CC.ref_lib()
    if lib is None:
        pytest.skip("oracle/_ref is not built")
    scene = k.Scene(CC.MIRRORS)
    device.upload(scene)
    qs = CC.queries()
    got = device.highlight_chains(qs)
    want = CC.ref_chains(lib, scene, qs)
    same = got["found"] == want["found"]
    both = (got["found"] == 1) & (want["found"] == 1)
    note(f"[chain] {len(qs)} chains: {int(want['found'].sum())} settle in the reference, "
         f"{int(got['found'].sum())} on the device, {int((~same).sum())} differ in settling")
    assert same.mean() >= 0.99
    assert both.sum() > len(qs) // 4
    dp = np.linalg.norm(got["point"][both] - want["point"][both], axis=1)
    da = np.linalg.norm(got["axis"][both] - want["axis"][both], axis=1)
    scale_s = np.abs(want["spread"][both]).max(axis=1)
    ds = np.abs(got["spread"][both] - want["spread"][both]).max(axis=1) / scale_s
    scale_m = np.abs(want["motion"][both]).reshape(-1, 6).max(axis=1) + 1e-12
    dm = np.abs(got["motion"][both] - want["motion"][both]).reshape(-1, 6).max(axis=1) / scale_m
    note(f"[chain] max |point| diff {dp.max():.2e}, axis {da.max():.2e}, spread rel {ds.max():.2e}, motion rel {dm.max():.2e}")
    assert dp.max() < 1e-7 and da.max() < 1e-7
    assert ds.max() < 1e-5 and dm.max() < 1e-5
    for kind in range(5):
        sel = qs["light"] == kind
        assert (both & sel).sum() > 0, f"no chain toward lamp kind {kind} settled"


def test_chain_queries_out_of_range_do_not_settle(device):
    scene = k.Scene(CC.MIRRORS)
    device.upload(scene)
    qs = CC.queries(n_starts=2)[:4].copy()
    qs["leaf"][0] = 1000
    qs["light"][1] = 99
    got = device.highlight_chains(qs)
    assert got["found"][0] == 0 and got["found"][1] == 0


def test_device_chains_match_the_reference_on_the_production_scene(device):
    """The production scene's own leaves (spheres, rounded boxes and tori
    under arbitrary turns) toward its four quad lamps: stands and seeds are
    camera hits (a stand on a matte or plastic surface, a seed on a metal
    leaf), 600 chains; the same chains settle as in the reference."""
    from tests import raysets
    lib = CC.ref_lib()
    if lib is None:
        pytest.skip("oracle/_ref is not built")
    scene = k.Scene(scenes_production())
    lower, size = scene.region(8)
    device.upload(scene)
    device.build_octree(lower, size, 8)
    rays = raysets.camera_rays(scene, 20000, 11)
    hits = device.extend(rays)
    mats = scene.materials()
    leaves = scene.leaves()
    surface = hits["kind"] == k.KDN_HIT_SURFACE
    leaf = hits["leaf"]
    kind = np.array([mats[leaves[l].material].kind if surface[i] and l < len(leaves) else -1 for i, l in enumerate(leaf)])
    metal = np.nonzero(kind == k.METAL)[0]
    stands = np.nonzero((kind == k.MATTE) | (kind == k.PLASTIC))[0]
    rng = np.random.default_rng(5)
    qs = np.zeros(600, k.CHAIN_QUERY_DTYPE)
    for i in range(len(qs)):
        m, s = metal[rng.integers(len(metal))], stands[rng.integers(len(stands))]
        n = hits["normal"][s].astype(np.float64)
        qs[i]["from"] = hits["point"][s] + n * 1e-3
        t = np.cross(n, [0, 0, 1] if abs(n[2]) < 0.9 else [1, 0, 0])
        t /= np.linalg.norm(t)
        qs[i]["sideways"] = [t, np.cross(n, t)]
        qs[i]["seed"] = hits["point"][m]
        qs[i]["width"] = 0.05
        qs[i]["leaf"] = leaf[m]
        qs[i]["light"] = rng.integers(4)
    got = device.highlight_chains(qs)
    want = CC.ref_chains(lib, scene, qs)
    same = got["found"] == want["found"]
    both = (got["found"] == 1) & (want["found"] == 1)
    dp = np.linalg.norm(got["point"][both] - want["point"][both], axis=1)
    note(f"[chain] production scene: {len(qs)} chains, {int(want['found'].sum())} settle in the reference, "
         f"{int(got['found'].sum())} on the device, {int((~same).sum())} differ; max |point| diff {dp.max() if len(dp) else 0:.2e}")
    assert same.mean() >= 0.99
    assert both.sum() >= 50
    assert dp.max() < 1e-7


def scenes_production():
    from kodanuki_amd import scenes
    return scenes.production(0, glass=False)
