"""The Sobol member on the device (plugins/sampler/sobol), through the C ABI:
- the device's eye-path numbers equal the oracle's to the bit: Sobol for the
  first 28 coordinates, hashed beyond them and without the member
  (sampler.cpp:15-30);
- a render drawing with it keeps the frame's mean and changes the frame
  (the reference's pairing check, target/tests/sampler.cpp:1103-1176);
- its per-pixel means agree with the FP64 oracle drawing the same numbers
  (paired, within 3 standard errors);
- at equal samples its error against a converged frame is lower than with
  hashed numbers, which is what the member is for."""
import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes
from oracle import sobol_ref as S
from tests import oracle_py as O
from tests.conftest import EMULATION

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)


STREAMS = [0x0123456789ABCDEF, 0xFEDCBA9876543210, 0x5555AAAA33331111, 0x9E3779B97F4A7C15]


@pytest.mark.parametrize("sobol", [True, False])
def test_device_numbers_equal_the_oracle(device, sobol):
    rng = np.random.default_rng(5)
    idx = np.concatenate([np.arange(4096), rng.integers(0, 1 << 32, 4096, dtype=np.int64)]).astype(np.uint32)
    L = O.lib()
    for stream in STREAMS:
        for coordinate in (0, 1, 5, 6, 7, 13, 27, 28, 40):
            got = device.sequence_numbers(stream, idx, coordinate, sobol=sobol)
            if sobol and coordinate < S.SOBOL_DIMENSIONS:
                want = S.sobol_number(np.full(len(idx), stream, np.uint64), idx.astype(np.uint64),
                                      np.full(len(idx), coordinate, np.uint64))
            else:
                want = np.array([L.or_hashed_number(stream, int(i), coordinate) for i in idx[:512]])
                got = got[:512]
            assert np.array_equal(got, want), (hex(stream), coordinate)


def _scene(device, text, depth):
    s = k.Scene(text)
    cam, rs = s.settings()
    lower, size = s.region(depth)
    device.upload(s)
    device.build_octree(lower, size, depth)
    return s, rs, lower, size


def _frame(device, W, H, spp, seed, surfaces, flags=0, first=0):
    device.render_pass(W, H, spp, first_sample=first, seed=seed, surfaces=surfaces, flags=flags)
    rgb, sq, n = device.accumulation(W, H)
    return rgb / n[..., None], sq / n[..., None], n


def test_sobol_keeps_the_mean_and_changes_the_frame(device):
    s, rs, lower, size = _scene(device, scenes.tiny(glass=False), 5)
    W, H, spp = 32, 24, 16
    means = {False: [], True: []}
    frames = {}
    for sobol in (False, True):
        for seed in range(1, 9):
            m, _, _ = _frame(device, W, H, spp, seed, 4, flags=k.PASS_SOBOL if sobol else 0)
            means[sobol].append(m.mean(axis=(0, 1)))
            if seed == 1:
                frames[sobol] = m
    a, b = np.array(means[False]), np.array(means[True])
    se = np.sqrt(a.var(axis=0, ddof=1) / len(a) + b.var(axis=0, ddof=1) / len(b))
    gap = np.max(np.abs(a.mean(axis=0) - b.mean(axis=0)) / np.maximum(se, 1e-12))
    note(f"[sobol] frame means hashed {a.mean(axis=0)} sobol {b.mean(axis=0)}, worst gap {gap:.2f} standard errors")
    assert gap < 4
    assert not np.array_equal(frames[False], frames[True])


def test_sobol_matches_fp64_oracle(device):
    text = scenes.tiny(glass=False)
    s, rs, lower, size = _scene(device, text, 5)
    W, H, spp = 24, 18, 64
    mg, sq, n = _frame(device, W, H, spp, rs.seed, 5, flags=k.PASS_SOBOL)
    var = np.maximum(sq - mg ** 2, 0)
    cpu = O.render64(s, O.SceneArrays(s), lower, size, W, H, spp, seed=rs.seed, surfaces=5, sobol=True)
    mc = cpu / spp
    z = np.abs(mg - mc) / (np.sqrt(var / spp) + 1e-4)
    note(f"[sobol] device vs FP64 oracle, the same Sobol numbers: within 3se {np.mean(z < 3):.4f}")
    assert np.mean(z < 3) > 0.99


@pytest.mark.skipif(bool(EMULATION), reason="a 1024-sample converged frame: device only")
def test_sobol_lowers_the_error(device):
    s, rs, lower, size = _scene(device, scenes.stopping_table(), 6)
    W, H = rs.width, rs.height
    truth, _, _ = _frame(device, W, H, 4096, 77, 6, first=1 << 24)
    err = {}
    for sobol in (False, True):
        e = []
        for seed in range(1, 5):
            m, _, _ = _frame(device, W, H, 16, seed, 6, flags=k.PASS_SOBOL if sobol else 0)
            e.append(np.mean((m - truth) ** 2 / (truth ** 2 + 0.01)))
        err[sobol] = float(np.mean(e))
    note(f"[sobol] relMSE at 16 spp against 4096 spp: hashed {err[False]:.3e}, sobol {err[True]:.3e} "
         f"({err[False] / err[True]:.2f}x)")
    assert err[True] < err[False]
