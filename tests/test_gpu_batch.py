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
