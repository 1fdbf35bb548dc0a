"""Renders of >= 8 M paths cycle their chunks through four streams (struct
Pipe): the image must be bit-identical to the same chunks run one
after another on one stream (KR_RENDER_ONE_STREAM), with and without a
radiance cache; batches, whose work spans their four streams, must
match their jobs rendered alone, with chunks cut per job and with chunks that
straddle jobs (KR_BATCH_MIX_JOBS)."""
import numpy as np
import pytest

from koishi_amd import kr, scenegen as sg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spheres():
    return sg.sphere_scene(width=1024, height=1024, spp=16)  # 16.8 M paths: four chunks, four streams


def _render(device, scene, flags=0, nrc=None, seed=5):
    gs = kr.GpuScene(device, scene.to_desc())
    r = kr.Render(gs, scene.camera_struct(), scene.params(seed=seed, flags=flags))
    if nrc is not None:
        r.attach_nrc(nrc)
    r.run(0, scene.spp)
    out = r.read()
    r.free()
    gs.free()
    return out


def test_two_streams_match_one(device, spheres):
    rgb2, val2 = _render(device, spheres)
    rgb1, val1 = _render(device, spheres, flags=kr.KR_RENDER_ONE_STREAM)
    assert np.array_equal(val2, val1) and np.array_equal(rgb2, rgb1)
    assert val1.sum() > 0.99 * spheres.width * spheres.height * spheres.spp


def test_two_streams_with_a_cache(device, spheres):
    # the first run's image depends on the cache's initial weights only (the
    # cache trains after the run), so both schedules give the same image
    pos = np.asarray(spheres.positions)
    out = []
    for flags in (0, kr.KR_RENDER_ONE_STREAM):
        nrc = kr.Nrc(device, pos.min(axis=0) - 10.0, pos.max(axis=0) + 10.0, seed=3)
        out.append(_render(device, spheres, flags=flags, nrc=nrc))
        nrc.free()
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][0], out[1][0])


def _batch_jobs():
    scenes = [sg.sphere_scene(), sg.mesh_scene(freq=24)]
    spec = [(0, 512, 512, 40, 5, 21), (1, 96, 40, 16, 8, 22), (0, 64, 48, 9, 5, 23), (1, 640, 360, 24, 8, 24)]
    return scenes, [(k, scenes[k].camera_struct(), kr.Params(w, h, spp, mb, seed, 0, 0))
                    for k, w, h, spp, mb, seed in spec]


@pytest.mark.parametrize("flags", [0, kr.KR_BATCH_MIX_JOBS])
def test_batch_on_two_streams_matches_jobs_alone(device, flags):
    scenes, jobs = _batch_jobs()
    ups = [kr.GpuScene(device, sc.to_desc(bvh_width=4)) for sc in scenes]
    ref = []
    for k, cam, p in jobs:
        r = kr.Render(ups[k], cam, p)
        r.run(0, p.spp)
        ref.append(r.read())
        r.free()
    for u in ups:
        u.free()
    total = sum(p.width * p.height * p.spp for _, _, p in jobs)
    assert total > 8 << 20
    P = 3_000_017
    b = kr.Batch(device, scenes, jobs, kr.BatchTuning(4, P, 32, flags, 0.0))
    for rep in range(2):
        b.reset()
        b.run()
        b.sync()
        info = b.info()
        if flags:
            assert info.chunks == -(-total // P)
        else:
            assert info.chunks == sum(-(-(p.width * p.height * p.spp) // P) for _, _, p in jobs)
        for j in range(len(jobs)):
            rgb, val = b.read(j)
            assert np.array_equal(val, ref[j][1]), (rep, j)
            assert np.allclose(rgb, ref[j][0], rtol=1e-12, atol=1e-12), (rep, j)
    b.free()


@pytest.mark.parametrize("which", ["mesh", "spheres"])
def test_tile_ordered_camera_queue_is_bit_identical(device, spheres, which, monkeypatch):
    """The camera rays of a chunk of whole samples enter the queue in 4 x 4
    pixel tiles of 4 samples (TileOrder): each path keeps its index, RNG
    streams and pixel, so the image equals the row-major queue's bit for bit
    (KR_CAMERA_TILE=0), on one stream and on four."""
    scene = sg.mesh_scene(width=256, height=128, spp=16) if which == "mesh" else spheres
    out = []
    for tile in ("0", "4x4x4", "8x8x1", "2x2x16"):
        monkeypatch.setenv("KR_CAMERA_TILE", tile)
        out.append(_render(device, scene))
    for rgb, val in out[1:]:
        assert np.array_equal(val, out[0][1]) and np.array_equal(rgb, out[0][0])


def _render_stats(device, scene, flags):
    gs = kr.GpuScene(device, scene.to_desc())
    r = kr.Render(gs, scene.camera_struct(), scene.params(seed=5, flags=flags))
    r.set_profiling(kr.KR_PROFILE_COUNTERS)
    r.run(0, scene.spp)
    r.sync()
    st = r.stats()
    counts = (st.closest_rays, st.node_visits, st.tri_tests)
    out = r.read()
    r.free()
    gs.free()
    return out, counts


@pytest.mark.parametrize("which", ["mesh", "spheres"])
def test_lane_refill_is_bit_identical(device, spheres, which):
    """Traversal waves that refill finished lanes from the queue
    (KR_RENDER_PERSISTENT_TRAVERSAL) walk each ray exactly as waves with a
    fixed share of rays do: the same node visits and triangle tests in total
    and the same hits, so the image is the same bit for bit; only the order
    in which hits enter the shading queues differs. The spheres-only scene has
    no BVH: every ray is done at its fetch."""
    scene = sg.mesh_scene(width=256, height=128, spp=16) if which == "mesh" else spheres
    flags = (0, kr.KR_RENDER_PERSISTENT_TRAVERSAL)
    runs = [_render_stats(device, scene, f) for f in flags]
    (rgb0, val0), c0 = runs[0]
    print(which, "closest rays, node visits, triangle tests:", c0, flush=True)
    assert c0[0] > 0 and (which == "spheres" or c0[1] > c0[0])
    for f, ((rgb, val), c) in zip(flags[1:], runs[1:]):
        assert c == c0, (f, c, c0)
        assert np.array_equal(val, val0) and np.array_equal(rgb, rgb0), f


@pytest.mark.parametrize("steps,rounds,stack", [("24", "2", "12"), ("1", "1", "12"), ("4", "4", "2"),
                                                ("8", "3", "64")])
def test_tail_rounds_are_bit_identical(device, monkeypatch, steps, rounds, stack):
    """Walks cut after a step budget and resumed by follow-up launches from
    the tail queues (k_trace_tail, k_shadow_tail) visit what one launch's
    walks visit: the same node visits and triangle tests in total and the
    same hits, so the image is the same bit for bit. The budgets force every
    path: queues that fill up (1 step: more walks outlast it than half a
    queue holds), stacks too deep for the queue (2 entries), several rounds,
    and tails from the camera rays on (KR_TAIL_FROM=0)."""
    scene = sg.mesh_scene(width=1024, height=512, spp=16)  # 8.4 M paths: four streams
    ref = _render_stats(device, scene, kr.KR_RENDER_NO_TAIL_ROUNDS)
    ref_plain = _render(device, scene, flags=kr.KR_RENDER_NO_TAIL_ROUNDS)
    monkeypatch.setenv("KR_TAIL_STEPS", steps)
    monkeypatch.setenv("KR_TAIL_ROUNDS", rounds)
    monkeypatch.setenv("KR_TAIL_STACK", stack)
    monkeypatch.setenv("KR_TAIL_FROM", "0")
    (rgb, val), c = _render_stats(device, scene, 0)
    (rgb0, val0), c0 = ref
    print("closest rays, node visits, triangle tests:", c0, c, flush=True)
    assert c == c0
    assert np.array_equal(val, val0) and np.array_equal(rgb, rgb0)
    rgb1, val1 = _render(device, scene)
    assert np.array_equal(val1, ref_plain[1]) and np.array_equal(rgb1, ref_plain[0])
    assert np.array_equal(val1, val0) and np.array_equal(rgb1, rgb0)
