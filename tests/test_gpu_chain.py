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
