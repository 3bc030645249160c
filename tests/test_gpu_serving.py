"""The serving loop (kdn_serve_*), through the C ABI.

A job's accumulators are the same to the bit alone and in a mix: its slices
(samples per round) are fixed by the job and the server's round size, its
paths draw numbers from (seed, pixel, sample) only, and its slices fold in
sample order into its own accumulators. The mixes here put jobs of three
scenes, ragged image sizes and sample counts into shared rounds, with jobs
arriving while others run, a budget that admits one job at a time, and a
round too small for a whole job. A one-slice job also equals the plain
pixel pass (kdn_render_pass) of the same image to the bit.
"""
import numpy as np
import pytest

import kodanuki_amd as k
from kodanuki_amd import scenes
from tests.conftest import EMULATION

pytestmark = pytest.mark.gpu

GLASS = bool(EMULATION)


def note(msg):
    print(msg, flush=True)


# (scene name, width, height, spp, seed, surfaces, depth)
JOBS = [
    ("plain0", 64, 48, 6, 11, 6, 6),
    ("tiny", 37, 21, 9, 5, 8, 5),
    ("plain3", 33, 17, 4, 2, 5, 6),
    ("plain0", 24, 40, 3, 7, 4, 6),   # the first job's scene again
    ("plain3", 96, 80, 12, 9, 5, 6),  # more paths than a round of 65536: slices of 8 and 4 samples
]


@pytest.fixture(scope="module")
def scene_set():
    return {"plain0": k.Scene(scenes.plain(0)), "tiny": k.Scene(scenes.tiny(glass=GLASS)), "plain3": k.Scene(scenes.plain(3))}


def _submit(server, scene_set, job):
    name, w, h, spp, seed, surfaces, depth = job
    return server.submit(scene_set[name], w, h, spp, seed=seed, surfaces=surfaces, depth=depth)


def _alone(device, scene_set, max_paths):
    out = []
    with k.Server(device, max_paths=max_paths) as server:
        for job in JOBS:
            j = _submit(server, scene_set, job)
            server.run([j])
            st = server.status(j)
            assert st.state == k.JOB_DONE, server.job_error(j)
            assert st.samples_done == job[3]
            out.append(server.result(j))
            server.release(j)
    return out


def _same(a, b):
    return all(x.tobytes() == y.tobytes() for x, y in zip(a, b))


@pytest.fixture(scope="module")
def alone(device, scene_set):
    return {mp: _alone(device, scene_set, mp) for mp in (1 << 16, 1 << 23)}


@pytest.mark.parametrize("max_paths", [1 << 16, 1 << 23])
def test_jobs_in_a_mix_match_jobs_alone(device, scene_set, alone, max_paths):
    with k.Server(device, max_paths=max_paths) as server:
        ids = [_submit(server, scene_set, job) for job in JOBS[:2]]
        st = server.step()
        note(f"[serve] first serving round: {st.jobs_in_round} jobs {st.slices_in_round} slices {st.paths} paths")
        # two more arrive while the first two run
        ids += [_submit(server, scene_set, job) for job in JOBS[2:]]
        rounds = 1 + server.run()
        for i, (j, job) in enumerate(zip(ids, JOBS)):
            status = server.status(j)
            assert status.state == k.JOB_DONE, server.job_error(j)
            assert status.submitted_ms <= status.admitted_ms <= status.finished_ms
            got = server.result(j)
            assert _same(got, alone[max_paths][i]), f"job {i} {job} differs in the mix"
            assert np.all(got[2] == job[3])
        note(f"[serve] max_paths {max_paths}: {len(JOBS)} jobs in {rounds} rounds, each equal to its render alone")


def _resident_sizes(device, scene_set):
    """Device bytes of the rounds' path state alone, and of each scene with
    a job of its own."""
    sizes = {}
    for job in JOBS:
        with k.Server(device, max_paths=1 << 16) as server:
            base = server.step().resident_bytes
            j = _submit(server, scene_set, job)
            server.run([j])
            sizes[job[0]] = max(sizes.get(job[0], 0), server.step().resident_bytes - base)
    return base, sizes


def test_budget_of_one_job_at_a_time(device, scene_set, alone):
    base, sizes = _resident_sizes(device, scene_set)
    # the largest scene with its job fits, no two scenes do
    budget = base + max(sizes.values()) + min(sizes.values()) // 2
    note(f"[serve] path state {base} bytes, scenes with a job {sizes}, budget {budget}")
    with k.Server(device, max_paths=1 << 16, budget_bytes=budget) as server:
        ids = [_submit(server, scene_set, job) for job in JOBS]
        rounds = 0
        while any(server.status(j).state in (k.JOB_WAITING, k.JOB_RUNNING) for j in ids):
            st = server.step()
            rounds += 1
            assert st.jobs_in_round <= 1 and st.running <= 1
            assert st.resident_bytes <= budget
        for i, j in enumerate(ids):
            assert server.status(j).state == k.JOB_DONE, server.job_error(j)
            assert _same(server.result(j), alone[1 << 16][i])
        note(f"[serve] one job at a time in {rounds} rounds, results equal to the jobs alone")


def test_job_beyond_the_budget_fails_alone(device, scene_set):
    # room for the path state of a round, not for a scene
    with k.Server(device, max_paths=1 << 16, budget_bytes=20 << 20) as server:
        j = _submit(server, scene_set, JOBS[0])
        server.step()
        assert server.status(j).state == k.JOB_FAILED
        assert "budget" in server.job_error(j)


def test_one_slice_job_equals_the_pixel_pass(device, scene_set):
    name, w, h, spp, seed, surfaces, depth = JOBS[0]
    s = scene_set[name]
    with k.Server(device, max_paths=1 << 23) as server:
        j = server.submit(s, w, h, spp, seed=seed, surfaces=surfaces, depth=depth)
        server.run()
        got = server.result(j)
    lower, size = s.region(depth)
    device.upload(s)
    device.build_octree(lower, size, depth)
    device.render_pass(w, h, spp, seed=seed, surfaces=surfaces)
    assert _same(got, device.accumulation(w, h))


def test_bad_jobs_are_refused(device, scene_set):
    with k.Server(device) as server:
        with pytest.raises(k.KodanukiError):
            server.submit(scene_set["plain0"], 0, 10, 1)
        with pytest.raises(k.KodanukiError):
            server.submit(scene_set["plain0"], 10, 10, 0)
        with pytest.raises(k.KodanukiError):
            server.submit(scene_set["plain0"], 10, 10, 1, depth=20)
    with pytest.raises(k.KodanukiError):
        k.Server(device, max_paths=1000)


@pytest.mark.skipif(bool(EMULATION), reason="the device's refusal; the CPU build runs glass")
def test_glass_job_fails_on_the_device(device):
    s = k.Scene(scenes.tiny())
    with k.Server(device) as server:
        j = server.submit(s, 16, 16, 1, depth=5)
        server.step()
        assert server.status(j).state == k.JOB_FAILED
        assert "glass" in server.job_error(j)


# guided jobs: (width, height, spp, seed) of the highlight room with the plain guide
GUIDED = [(64, 48, 8, 3), (96, 72, 12, 4)]


@pytest.fixture(scope="module")
def room():
    return k.Scene(scenes.highlight_room())


def test_guided_one_slice_job_equals_the_guided_pixel_pass(device, room):
    """A served job with the plain guide: its survey at admission gives
    the guides kdn_highlight_guides gives, and its one slice the pixel
    pass of pass flag bit 5, to the bit."""
    w, h, spp, seed = GUIDED[0]
    with k.Server(device, max_paths=1 << 23) as server:
        j = server.submit(room, w, h, spp, seed=seed, surfaces=6, depth=6, guided=True)
        server.run()
        assert server.status(j).state == k.JOB_DONE, server.job_error(j)
        got = server.result(j)
    lower, size = room.region(6)
    device.upload(room)
    device.build_octree(lower, size, 6)
    st = device.highlight_guides(w, h)
    assert st.guided_pixels > 0
    device.render_pass(w, h, spp, seed=seed, surfaces=6, flags=k.PASS_HIGHLIGHT)
    want = device.accumulation(w, h)
    assert _same(got, want)
    device.render_pass(w, h, spp, seed=seed, surfaces=6)
    assert not _same(got, device.accumulation(w, h))  # the guide changed the samples


def test_guided_jobs_in_a_mix_match_jobs_alone(device, room, scene_set):
    jobs = [("guided", *GUIDED[0]), ("plain", *JOBS[0][1:5]), ("guided", *GUIDED[1])]

    def submit(server, job):
        kind, w, h, spp, seed = job
        if kind == "guided":
            return server.submit(room, w, h, spp, seed=seed, surfaces=6, depth=6, guided=True)
        return server.submit(scene_set["plain0"], w, h, spp, seed=seed, surfaces=6, depth=6)

    alone = []
    with k.Server(device, max_paths=1 << 16) as server:
        for job in jobs:
            j = submit(server, job)
            server.run([j])
            assert server.status(j).state == k.JOB_DONE, server.job_error(j)
            alone.append(server.result(j))
    with k.Server(device, max_paths=1 << 16) as server:
        ids = [submit(server, jobs[0]), submit(server, jobs[1])]
        server.step()
        ids.append(submit(server, jobs[2]))
        server.run()
        for i, j in enumerate(ids):
            assert server.status(j).state == k.JOB_DONE, server.job_error(j)
            assert _same(server.result(j), alone[i]), i


@pytest.mark.skipif(bool(EMULATION), reason="scenes of up to 4096 primitives at depth 8: device only")
def test_pool_stream_jobs_match_jobs_alone(device):
    """The serving stream's own scene pool (eight seeded scenes of 32 to 4096
    primitives, octree depths 6 to 8) with a seeded Poisson arrival order at
    rates 2, 8 and 32 jobs/s: arrivals are mapped to server steps (a round
    takes about a tenth of a second at these sizes), so the mix is fixed by
    the seed; every job's accumulators equal the same job served alone."""
    import random
    pool = [(k.Scene(text), depth) for text, depth in scenes.serving_pool(glass=GLASS)]
    sizes = [(64, 64), (128, 96), (160, 90), (256, 256)]
    spps = [1, 4, 16]
    rng = random.Random(5)
    for rate in (2.0, 8.0, 32.0):
        t, stream = 0.0, []
        for _ in range(10):
            t += rng.expovariate(rate)
            w, h = rng.choice(sizes)
            stream.append((int(t / 0.1), rng.randrange(len(pool)), w, h, rng.choice(spps), rng.randrange(1, 100)))
        alone = []
        with k.Server(device, max_paths=1 << 18) as server:
            for _, si, w, h, spp, seed in stream:
                scene, depth = pool[si]
                j = server.submit(scene, w, h, spp, seed=seed, surfaces=6, depth=depth)
                server.run([j])
                assert server.status(j).state == k.JOB_DONE, server.job_error(j)
                alone.append(server.result(j))
                server.release(j)
        with k.Server(device, max_paths=1 << 18) as server:
            ids, pending, steps, most = [], list(stream), 0, 0
            while pending or any(server.status(j).state not in (k.JOB_DONE, k.JOB_FAILED) for j in ids):
                while pending and pending[0][0] <= steps:
                    _, si, w, h, spp, seed = pending.pop(0)
                    scene, depth = pool[si]
                    ids.append(server.submit(scene, w, h, spp, seed=seed, surfaces=6, depth=depth))
                st = server.step()
                most = max(most, st.jobs_in_round)
                steps += 1
            for i, j in enumerate(ids):
                assert server.status(j).state == k.JOB_DONE, server.job_error(j)
                assert _same(server.result(j), alone[i]), (rate, i, stream[i])
        note(f"[serving] pool stream at {rate} jobs/s: {len(stream)} jobs over {steps} steps, up to {most} jobs in a round, "
             f"each equal alone and in the mix")
