"""Batched renders (kr_batch_*): many uneven jobs in one wavefront, replayed
from one captured HIP graph. Every job must reproduce the same job rendered
alone (same scene at the same BVH width, camera, params): a path's RNG
streams, lens sample and pixel depend only on its job-local index, so every
path's radiance is bit-identical; the valid counts must match exactly and
the FP64 sums up to the association of the per-chunk partial sums (the
batch's chunks cut the jobs elsewhere), i.e. to 1e-12 relative."""
import numpy as np
import pytest

from koishi_amd import kr, scenegen as sg

pytestmark = pytest.mark.gpu


def _jobs():
    scenes = [sg.sphere_scene(), sg.mesh_scene(freq=24), sg.room_scene(freq=12),
              sg.large_scene(instances=3, freq=10)]
    # (scene, width, height, spp, max_bounce, seed): uneven sizes, depths and seeds
    spec = [(0, 64, 48, 9, 5, 11), (1, 96, 40, 16, 8, 12), (2, 40, 30, 4, 3, 13), (3, 128, 72, 2, 8, 14),
            (1, 17, 13, 5, 8, 15), (0, 33, 21, 7, 2, 16)]
    jobs = []
    for k, w, h, spp, mb, seed in spec:
        s = scenes[k]
        p = kr.Params(w, h, spp, mb, seed, 0, 0)
        jobs.append((k, s.camera_struct(), p))
    return scenes, jobs


def _alone(device, scenes, jobs, width):
    out = []
    ups = [kr.GpuScene(device, sc.to_desc(bvh_width=width)) for sc in scenes]
    for k, cam, p in jobs:
        r = kr.Render(ups[k], cam, p)
        r.run(0, p.spp)
        out.append(r.read())
        r.free()
    for u in ups:
        u.free()
    return out


def test_batch_matches_individual_renders(device):
    scenes, jobs = _jobs()
    ref = _alone(device, scenes, jobs, 4)
    # small chunks that straddle jobs: every chunk boundary and many job
    # boundaries fall inside chunks, which run the batch kernels
    tun = kr.BatchTuning(4, 20011, 8, kr.KR_BATCH_MIX_JOBS, 0.0)
    b = kr.Batch(device, scenes, jobs, tun)
    b.run()
    b.sync()
    info = b.info()
    total = sum(p.width * p.height * p.spp for _, _, p in jobs)
    assert info.total_paths == total and info.num_jobs == len(jobs) and info.num_scenes == len(scenes)
    assert info.chunks == -(-total // 20011) and info.graph_nodes > info.chunks
    for j in range(len(jobs)):
        rgb, val = b.read(j)
        assert np.array_equal(val, ref[j][1]), j
        assert np.allclose(rgb, ref[j][0], rtol=1e-12, atol=1e-12), j
    # graph replays (after a reset) reproduce the first launch exactly
    first = [b.read(j) for j in range(len(jobs))]
    for _ in range(2):
        b.reset()
        b.run()
        b.sync()
        for j in range(len(jobs)):
            rgb, val = b.read(j)
            assert np.array_equal(rgb, first[j][0]) and np.array_equal(val, first[j][1])
    assert b.info().seconds > 0
    b.free()


def test_batch_autotune_picks_a_valid_tuning(device):
    scenes, jobs = _jobs()
    b = kr.Batch(device, scenes, jobs)
    t = b.autotune(trial_spp=2)
    assert t.bvh_width in (2, 4, 8)
    assert t.chunk_paths in (16 << 20, 32 << 20) or t.chunk_paths <= sum(
        p.width * p.height * p.spp for _, _, p in jobs)
    assert t.blocks_per_cu in (32, 64) and t.paths_per_second > 0
    info = b.info()
    assert info.tuning.bvh_width == t.bvh_width
    # accumulators were cleared after the trials; the tuned batch still
    # matches the jobs rendered alone at the chosen width
    rgb, val = b.read(0)
    assert not val.any()
    b.run()
    b.sync()
    ref = _alone(device, scenes, jobs, t.bvh_width)
    for j in range(len(jobs)):
        rgb, val = b.read(j)
        assert np.array_equal(val, ref[j][1]) and np.allclose(rgb, ref[j][0], rtol=1e-12, atol=1e-12), j
    b.free()


def test_batch_rejects_bad_input(device):
    scenes, jobs = _jobs()
    with pytest.raises(kr.KrError):
        kr.Batch(device, scenes, [(len(scenes), jobs[0][1], jobs[0][2])])
    with pytest.raises(kr.KrError):
        kr.Batch(device, scenes, jobs, kr.BatchTuning(3, 0, 8, 0, 0.0))
    with pytest.raises(kr.KrError):
        kr.Batch(device, scenes, jobs, kr.BatchTuning(4, 0, 65, 0, 0.0))
    b = kr.Batch(device, scenes[:1], jobs[:1])
    with pytest.raises(kr.KrError):
        b.read(1)
    b.free()


def _cached_jobs():
    scenes = [sg.room_scene(freq=12), sg.mesh_scene(freq=24), sg.sphere_scene()]
    spec = [(0, 256, 192, 16, 8, 31), (1, 160, 80, 16, 8, 32), (0, 128, 96, 8, 8, 33), (2, 96, 96, 9, 5, 34)]
    return scenes, [(k, scenes[k].camera_struct(), kr.Params(w, h, spp, mb, seed, 0, 0))
                    for k, w, h, spp, mb, seed in spec]


def _box(scene):
    pos = np.asarray(scene.positions, np.float64)
    if len(scene.spheres):
        c = np.array([s[0] for s in scene.spheres])
        r = np.array([s[1] for s in scene.spheres])[:, None]
        pos = np.concatenate([pos, c - np.minimum(r, 20), c + np.minimum(r, 20)])
    return pos.min(axis=0) - 0.01, pos.max(axis=0) + 0.01


def _params(seed):
    rng = np.random.default_rng(seed)
    return (rng.normal(0, 0.08, kr.NRC_PARAMS) + 0.02).astype(np.float32)


@pytest.mark.parametrize("mode", [0, kr.NRC_INDIRECT])
def test_cached_batch_matches_cached_jobs_alone(device, mode):
    """A batch whose jobs 0 and 2 (one scene) share a radiance cache and job
    3 has its own: with the caches frozen at fixed weights, every job equals
    the same job rendered alone with the same cache (identical valid counts,
    FP64 sums to 1e-12); job 1 has no cache and equals its plain render."""
    scenes, jobs = _cached_jobs()
    caches = {}
    for k in (0, 2):
        n = kr.Nrc(device, *_box(scenes[k]), train_log2=3, flags=kr.NRC_FROZEN | mode)
        n.set_params(_params(40 + k))
        caches[k] = n
    b = kr.Batch(device, scenes, jobs)
    for j, (k, _, _) in enumerate(jobs):
        if k in caches:
            b.attach_nrc(j, caches[k])
    b.run()
    b.sync()
    got = [b.read(j) for j in range(len(jobs))]
    b.free()
    for j, (k, cam, p) in enumerate(jobs):
        gs = kr.GpuScene(device, scenes[k].to_desc(bvh_width=4))
        r = kr.Render(gs, cam, p)
        if k in caches:
            r.attach_nrc(caches[k])
        r.run(0, p.spp)
        rgb, val = r.read()
        q = r.stats().cache_queries
        r.free()
        gs.free()
        assert (q > 0) == (k in caches), (j, q)
        assert np.array_equal(val, got[j][1]), j
        assert np.allclose(rgb, got[j][0], rtol=1e-12, atol=1e-12), j
    for n in caches.values():
        n.free()


def test_cached_batch_trains_its_caches(device):
    """Trainable caches on a batch: one kr_batch_run gathers the training
    samples of all the cache's jobs (the same multiset as rendering those jobs
    alone: training paths never query the cache) and then takes
    steps_per_frame Adam steps per cache, once."""
    scenes, jobs = _cached_jobs()

    def fresh(k):
        # 1 path in 64 trains: the jobs' records stay below the 2^18-entry
        # pool, which therefore holds all of them, whatever their order
        n = kr.Nrc(device, *_box(scenes[k]), train_log2=6, steps_per_frame=3, batch=4096)
        n.set_params(_params(50 + k))
        return n
    cb = {k: fresh(k) for k in (0, 2)}
    b = kr.Batch(device, scenes, jobs)
    for j, (k, _, _) in enumerate(jobs):
        if k in cb:
            b.attach_nrc(j, cb[k])
    b.run()
    b.sync()
    b.free()
    for k, n in cb.items():
        trained, _ = n.stats()
        assert trained == 3 * 4096, (k, trained)
    alone = {k: fresh(k) for k in (0, 2)}
    for k, cam, p in jobs:
        if k not in alone:
            continue
        gs = kr.GpuScene(device, scenes[k].to_desc(bvh_width=4))
        r = kr.Render(gs, cam, p)
        r.attach_nrc(alone[k])
        r.run(0, p.spp)
        r.sync()
        r.free()
        gs.free()
    for k in (0, 2):
        raw_b, tgt_b = cb[k].read_pool(1 << 18)
        raw_a, tgt_a = alone[k].read_pool(1 << 18)
        assert raw_b.shape[0] > 500 and raw_b.shape == raw_a.shape, (k, raw_b.shape, raw_a.shape)
        rows_b = np.concatenate([raw_b, tgt_b], axis=1)
        rows_a = np.concatenate([raw_a, tgt_a], axis=1)
        ob = np.lexsort(rows_b.T[::-1])
        oa = np.lexsort(rows_a.T[::-1])
        assert np.array_equal(rows_b[ob], rows_a[oa]), k
    for n in list(cb.values()) + list(alone.values()):
        n.free()


def test_cached_batch_rejects_mixed_chunks_and_tuning(device):
    scenes, jobs = _cached_jobs()
    n = kr.Nrc(device, *_box(scenes[0]))
    b = kr.Batch(device, scenes, jobs, kr.BatchTuning(4, 20011, 8, kr.KR_BATCH_MIX_JOBS, 0.0))
    with pytest.raises(kr.KrError):
        b.attach_nrc(0, n)
    b.free()
    b = kr.Batch(device, scenes, jobs)
    b.attach_nrc(0, n)
    with pytest.raises(kr.KrError):
        b.autotune(trial_spp=1)
    b.attach_nrc(0, None)
    b.free()
    n.free()
