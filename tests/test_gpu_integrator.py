"""GPU parity of the wavefront integrator, through the C ABI.

- furnace: a white matte, a white metal (with its energy compensation) and
  a dispersive glass under a uniform white sky converge to radiance one,
  the closed form of a lossless scene;
- determinism: two progressive passes of 8 samples are one pass of 16 to
  the bit (counter-based numbers, folds in sample order);
- the FP64 oracle: on a small image of a scene with matte, metal, glass and
  a quad lamp, per-pixel means agree within 3 standard errors and the
  block means within a relative MSE of 1e-3.
"""
import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes
from tests import oracle_py as O
from tests.conftest import EMULATION

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)


def furnace(material):
    return f"""import "standard.sdf";
render {{ width = 32; height = 32; seed = 3; samples = 64; surfaces = 64; far = 2; }}
camera {{ eye = [0, -3, 0]; target = [0, 0, 0]; up = [0, 0, 1]; fov = 0.6; }}
sky gradient {{ zenith = [1, 1, 1]; horizon = [1, 1, 1]; ground = [1, 1, 1]; }}
let m = {material};
paint(m) sphere(0.5);
paint(m) at([0.3, 0.3, 0.3]) box([0.4, 0.4, 0.4], 0.05);
"""


def _render(device, text, depth, spp, w=None, h=None, surfaces=None, seed=None, keep=False, first=0, min_depth=4, flags=0):
    s = k.Scene(text)
    cam, rs = s.settings()
    w, h = w or rs.width, h or rs.height
    lower, size = s.region(depth)
    device.upload(s)
    device.build_octree(lower, size, depth)
    st = device.render_pass(w, h, spp, first_sample=first, seed=seed or rs.seed, surfaces=surfaces or rs.surfaces,
                            min_depth=min_depth, keep=keep, flags=flags)
    rgb, sq, n = device.accumulation(w, h)
    return s, lower, size, rgb, sq, n, st


# Glass scattering is held off the device (kdn_render_pass refuses scenes
# with glass there): two renders of the glass furnace faulted the device and
# no cause is shown. The bounds-checking CPU build of the library runs the
# glass cases (KDN_EMULATION); on the device the scenes paint the glass as a
# white matte.
GLASS = bool(EMULATION)
_GLASS_HELD = pytest.mark.skipif(not EMULATION, reason="glass scattering is held off the device: its furnace render "
                                                       "faulted the device twice, cause not found")


@pytest.mark.parametrize("material", ["matte([1, 1, 1])", "metal([1, 1, 1], roughness = 0.6)",
                                      pytest.param("glass(refraction = 1.5, dispersion = 0.008)", marks=_GLASS_HELD)])
def test_white_furnace(device, material):
    s, lower, size, rgb, sq, n, st = _render(device, furnace(material), 5, 512)
    mean = rgb / n[..., None]
    note(f"[furnace] {material}: mean {mean.mean(axis=(0, 1))} min {mean.min():.4f} max {mean.max():.4f} "
         f"segments {st.segments} {st.ms:.1f} ms")
    # lossless: every pixel converges to one (the metal's compensation is a
    # tabled fit, within a few percent of energy conservation)
    tol = 0.03 if "metal" in material else 0.01
    assert abs(mean.mean() - 1.0) < tol
    if "matte" in material:
        var = sq / n[..., None] - mean ** 2
        se = np.sqrt(np.maximum(var, 1e-12) / n[..., None])
        assert np.mean(np.abs(mean - 1.0) < 4 * se + 1e-6) > 0.99


@pytest.mark.skipif(bool(EMULATION), reason="the device's refusal; the CPU build runs glass")
def test_device_refuses_glass(device):
    s = k.Scene(scenes.tiny())
    lower, size = s.region(5)
    device.upload(s)
    device.build_octree(lower, size, 5)
    with pytest.raises(k.KodanukiError, match="glass scattering is held off the device"):
        device.render_pass(24, 18, 1)


def test_progressive_passes_equal_one_pass(device):
    text = scenes.tiny(glass=GLASS)
    _, _, _, a, _, na, _ = _render(device, text, 5, 16)
    _render(device, text, 5, 8)
    _, _, _, b, _, nb, _ = _render(device, text, 5, 8, first=8, keep=True)
    assert np.array_equal(na, nb)
    assert np.allclose(a, b, rtol=0, atol=1e-9)


def test_finishing_launch_changes_nothing(device):
    """The paths finished in one launch get the same estimates, to the bit,
    as bounce-by-bounce wavefront launches (same per-path arithmetic, folds
    in sample order), and the same segment and shadow counts."""
    text = scenes.tiny(glass=GLASS)
    W, H, spp = 24, 18, 512
    _, _, _, a, sa, na, st_a = _render(device, text, 5, spp, w=W, h=H)
    _, _, _, b, sb, nb, st_b = _render(device, text, 5, spp, w=W, h=H, flags=k.PASS_NO_FINISH)
    note(f"[finish] segments {st_a.segments} / {st_b.segments}, finish {st_a.finish_ms:.1f} ms")
    assert np.array_equal(na, nb)
    assert a.tobytes() == b.tobytes() and sa.tobytes() == sb.tobytes()
    assert (st_a.segments, st_a.shadow_rays) == (st_b.segments, st_b.shadow_rays)


def test_matches_fp64_oracle(device):
    text = scenes.tiny(glass=GLASS)
    W, H = 24, 18
    spp_gpu, spp_cpu = 8192, 256
    s, lower, size, rgb, sq, n, st = _render(device, text, 5, spp_gpu, w=W, h=H, surfaces=5)
    note(f"[oracle] gpu {W}x{H}x{spp_gpu}: {st.ms:.1f} ms, {st.segments} segments")
    mg = rgb / n[..., None]
    vg = np.maximum(sq / n[..., None] - mg ** 2, 0)
    A = O.SceneArrays(s)
    cpu = O.render64(s, A, lower, size, W, H, spp_cpu, seed=s.settings()[1].seed, surfaces=5)
    mc = cpu / spp_cpu
    # the oracle's variance from the GPU's per-sample variance (same estimator)
    se = np.sqrt(vg / spp_cpu + vg / spp_gpu) + 1e-4
    z = np.abs(mg - mc) / se
    note(f"[oracle] mean gpu {mg.mean(axis=(0, 1))} cpu {mc.mean(axis=(0, 1))} within 3se {np.mean(z < 3):.4f}")
    assert np.mean(z < 3) > 0.97
    # block means (3x3) against the oracle: relMSE = mean (a-b)^2/(b^2+0.01)
    def blocks(img):
        return img.reshape(H // 3, 3, W // 3, 3, 3).mean(axis=(1, 3))
    bg, bc = blocks(mg), blocks(mc)
    rel = np.mean((bg - bc) ** 2 / (bc ** 2 + 0.01))
    noise = np.mean(blocks(vg / spp_cpu / 9) / (bc ** 2 + 0.01))
    note(f"[oracle] block relMSE {rel:.2e} (oracle noise {noise:.2e})")
    assert rel < 1e-3 + 3 * noise


@pytest.mark.skipif(bool(EMULATION), reason="a million device samples and a 256x256 oracle render: device only")
def test_config0_matches_fp64_oracle(device):
    """The plain scene at its own size (32 primitives smooth-merged over a
    plane, one quad lamp, 256x256, 16 spp, 6 bounces, octree depth 6)
    against the FP64 oracle's render of the same 16 samples of every pixel
    (the same counter-based numbers: paired estimates), and against an
    independent 256-spp device render (per-pixel means within 3 standard
    errors; 8x8 block means within relMSE 1e-3 of each other beyond the
    oracle's own noise)."""
    text = scenes.plain(0)
    s = k.Scene(text)
    cam, rs = s.settings()
    W, H, spp = rs.width, rs.height, rs.samples
    assert (W, H, spp, rs.surfaces) == (256, 256, 16, 7)
    _, lower, size, g16, _, n16, _ = _render(device, text, 6, spp, w=W, h=H, surfaces=rs.surfaces, seed=rs.seed)
    mg16 = g16 / n16[..., None]
    _, _, _, ghi, sqhi, nhi, st = _render(device, text, 6, 256, w=W, h=H, surfaces=rs.surfaces, seed=rs.seed, first=1 << 20)
    mhi = ghi / nhi[..., None]
    var = np.maximum(sqhi / nhi[..., None] - mhi ** 2, 0)
    A = O.SceneArrays(s)
    cpu = O.render64(s, A, lower, size, W, H, spp, seed=rs.seed, surfaces=rs.surfaces)
    mc = cpu / spp
    se16 = np.sqrt(var / spp) + 1e-4
    paired = np.abs(mg16 - mc) / se16
    unpaired = np.abs(mhi - mc) / (np.sqrt(var / spp + var / 256) + 1e-4)
    rel16 = np.mean((mg16 - mc) ** 2 / (mc ** 2 + 0.01))
    note(f"[config0] gpu 16 spp vs oracle 16 spp (same numbers): within 3se {np.mean(paired < 3):.5f}, relMSE {rel16:.2e}; "
         f"gpu 256 spp (other numbers) vs oracle: within 3se {np.mean(unpaired < 3):.5f}; means gpu {mhi.mean(axis=(0, 1))} "
         f"oracle {mc.mean(axis=(0, 1))}")
    assert np.mean(paired < 3) > 0.99
    assert np.mean(unpaired < 3) > 0.98

    def blocks(img):
        return img.reshape(H // 8, 8, W // 8, 8, 3).mean(axis=(1, 3))

    bg, bc = blocks(mhi), blocks(mc)
    rel = np.mean((bg - bc) ** 2 / (bc ** 2 + 0.01))
    noise = np.mean(blocks(var / spp / 64) / (bc ** 2 + 0.01))
    note(f"[config0] 8x8 block relMSE {rel:.2e} (oracle noise {noise:.2e})")
    assert rel < 1e-3 + 3 * noise
