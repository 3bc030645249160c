"""Served jobs with the learned guide (kdn_job_desc flags bit 1), through the
C ABI.

Each such job owns a guide over its scene's region, made from the job's
seed, trained after every round on the records of its own paths only, and
the job takes one slice a round: its slice k is steered by the guide as
trained on its slices before k, whatever else shares the rounds, and its
training sums run in a fixed order. So its accumulators are the same to
the bit alone and in a mix, beside plain jobs, plain-guided jobs and other
learned jobs, as every served job's are.
"""
import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes
from tests.conftest import EMULATION

pytestmark = pytest.mark.gpu


def note(msg):
    print(msg, flush=True)


# (kind, width, height, spp, seed): at 65536 paths a round, 9 samples a
# slice at 96x72, 12 at 80x64 and 256 at 16x16, so the learned jobs take 2
# to 3 slices (a whole slice fills a round of 65536); the first job's last
# slice (8 samples) shares its round, so a round with a learned job
# finishes several jobs' paths apart past the guided depths
JOBS = [
    ("learned", 16, 16, 264, 6),
    ("plain", 64, 48, 6, 11),
    ("learned", 96, 72, 18 if EMULATION else 24, 3),
    ("learned", 80, 64, 24, 4),
    ("guided", 64, 48, 8, 5),
]
# rounds of 2^23 paths hold several 2^21-path slices, and a learned job's
# slices are up to 2^22 paths there (606 samples at 96x72, 819 at 80x64):
# learned jobs of 2 slices share rounds with each other and with plain and
# plain-guided jobs
SHARED = [
    ("learned", 96, 72, 700, 3),
    ("plain", 64, 48, 6, 11),
    ("learned", 80, 64, 900, 4),
    ("guided", 64, 48, 8, 5),
]


@pytest.fixture(scope="module")
def scene_set():
    return {"window": k.Scene(scenes.window_room(96, 72)), "plain0": k.Scene(scenes.plain(0)),
            "highlight": k.Scene(scenes.highlight_room())}


def _submit(server, scene_set, job, learned=None):
    kind, w, h, spp, seed = job
    if kind == "plain":
        return server.submit(scene_set["plain0"], w, h, spp, seed=seed, surfaces=6, depth=6)
    if kind == "guided":
        return server.submit(scene_set["highlight"], w, h, spp, seed=seed, surfaces=6, depth=6, guided=True)
    return server.submit(scene_set["window"], w, h, spp, seed=seed, surfaces=6, depth=6,
                         learned=True if learned is None else learned)


def _same(a, b):
    return all(x.tobytes() == y.tobytes() for x, y in zip(a, b))


def _served_alone(device, scene_set, job, learned=None, max_paths=1 << 16):
    """The job served alone: its accumulators and each round's statistics."""
    with k.Server(device, max_paths=max_paths) as server:
        j = _submit(server, scene_set, job, learned)
        rounds = []
        while server.status(j).state in (k.JOB_WAITING, k.JOB_RUNNING):
            rounds.append(server.step())
        assert server.status(j).state == k.JOB_DONE, server.job_error(j)
        return server.result(j), rounds


def test_learned_job_trains_between_its_slices(device, scene_set):
    """Alone, a learned job runs one slice a round; the guide trains after
    each but the last, and its later slices sample with it: the image
    differs from the same job without the guide, stays finite and near its
    mean."""
    job = JOBS[2]
    learned, rounds = _served_alone(device, scene_set, job)
    plain, _ = _served_alone(device, scene_set, job, learned=False)
    spp = job[3]
    steps = [r.pass_.guide_steps for r in rounds if r.slices_in_round]
    samples = [r.pass_.guide_samples for r in rounds if r.slices_in_round]
    note(f"[serve-learned] {len(steps)} rounds, guide steps {steps}, samples {samples}")
    assert len(steps) == -(-spp // 9) and all(r.slices_in_round <= 1 for r in rounds)
    assert all(n > 0 for n in steps[:-1]) and all(n > 0 for n in samples[:-1])
    assert steps[-1] == 0 and samples[-1] == 0  # nothing samples with a guide trained after the last slice
    assert np.all(learned[2] == spp) and np.all(np.isfinite(learned[0]))
    assert not _same(learned, plain)  # the later slices sampled with the guide
    mean_l = learned[0].sum() / learned[2].sum()
    mean_p = plain[0].sum() / plain[2].sum()
    note(f"[serve-learned] image means: learned {mean_l:.4f}, plain {mean_p:.4f}")
    assert abs(mean_l / mean_p - 1) < 0.25


@pytest.mark.parametrize("mix", ["small", "shared"])
def test_learned_jobs_in_a_mix_match_jobs_alone(device, scene_set, mix):
    if mix == "shared" and EMULATION:
        pytest.skip("rounds of millions of paths: device only")
    jobs, max_paths = (JOBS, 1 << 16) if mix == "small" else (SHARED, 1 << 23)
    alone = [_served_alone(device, scene_set, job, max_paths=max_paths)[0] for job in jobs]
    with k.Server(device, max_paths=max_paths) as server:
        ids = [_submit(server, scene_set, job) for job in jobs[:2]]
        st = server.step()
        note(f"[serve-learned] first round: {st.jobs_in_round} jobs, {st.slices_in_round} slices")
        ids += [_submit(server, scene_set, job) for job in jobs[2:]]
        most = st.jobs_in_round
        while any(server.status(j).state in (k.JOB_WAITING, k.JOB_RUNNING) for j in ids):
            most = max(most, server.step().jobs_in_round)
        for i, (j, job) in enumerate(zip(ids, jobs)):
            assert server.status(j).state == k.JOB_DONE, server.job_error(j)
            assert _same(server.result(j), alone[i]), f"job {i} {job} differs in the mix"
    note(f"[serve-learned] {mix}: {len(jobs)} jobs, up to {most} a round, each equal to its render alone")
    assert most >= (2 if mix == "small" else 3)


def test_learned_job_of_one_slice_is_the_plain_job(device, scene_set):
    """A learned job whose samples fit one slice never samples with a
    guide: it gets none, and equals the same job without the flag."""
    job = ("learned", 32, 24, 4, 9)
    learned, rounds = _served_alone(device, scene_set, job)
    plain, _ = _served_alone(device, scene_set, job, learned=False)
    assert _same(learned, plain)
    assert all(r.pass_.guide_steps == 0 for r in rounds)


def test_learned_and_plain_guide_together_are_refused(device, scene_set):
    with k.Server(device) as server:
        with pytest.raises(k.KodanukiError):
            server.submit(scene_set["window"], 16, 16, 1, depth=6, guided=True, learned=True)


@pytest.mark.skipif(bool(EMULATION), reason="thousands of samples per pixel: device only")
def test_served_learned_job_is_unbiased(device):
    """A served learned job's block means agree with a 512-spp plain render
    of other samples within 4 standard errors (97 % of blocks)."""
    text = scenes.window_room(128, 96)
    scene = k.Scene(text)
    with k.Server(device, max_paths=1 << 18) as server:
        j = server.submit(scene, 128, 96, 64, seed=7, surfaces=6, depth=6, learned=True)
        rounds = 0
        while server.status(j).state in (k.JOB_WAITING, k.JOB_RUNNING):
            server.step()
            rounds += 1
        assert server.status(j).state == k.JOB_DONE, server.job_error(j)
        rgb, sq, n = server.result(j)
    _, rs = scene.settings()
    lower, size = scene.region(6)
    device.upload(scene)
    device.build_octree(lower, size, 6)
    device.render_pass(128, 96, 512, first_sample=1 << 20, seed=7, surfaces=6)
    ref_rgb, _, ref_n = device.accumulation(128, 96)
    ref = ref_rgb / ref_n[..., None]
    img = rgb / n[..., None]
    var = sq / n[..., None] - img ** 2
    blocks = img.reshape(12, 8, 16, 8, 3).mean(axis=(1, 3))
    ref_blocks = ref.reshape(12, 8, 16, 8, 3).mean(axis=(1, 3))
    se = np.sqrt(np.maximum(var, 0).reshape(12, 8, 16, 8, 3).sum(axis=(1, 3)) / (64 * 64 * 64))
    z = np.abs(blocks - ref_blocks) / (se + 1e-6)
    note(f"[serve-learned] 128x96 at 64 spp in {rounds} rounds: block |z| max {z.max():.2f}, mean {z.mean():.2f}; "
         f"image mean {img.mean():.4f} against {ref.mean():.4f}")
    assert np.mean(z < 4) > 0.97
    assert abs(img.mean() / ref.mean() - 1) < 0.05
