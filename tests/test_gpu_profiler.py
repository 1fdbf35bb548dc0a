"""The profiler (core::Profiler over RadianceKernel::Slice) on the GPU,
through kr_render_set_profiler / kr_render_read_slices, against the FP64
oracle's slices of the same seeded paths; the profiled area restriction; and
the rejected combinations."""
import numpy as np
import pytest

import oracle_lib as ol
from koishi_amd import kr, scenegen as sg

pytestmark = pytest.mark.gpu

AREA = (5, 4, 17, 13)  # 12 x 9 pixels of a 24 x 18 frame
DEPTH = 6


@pytest.fixture(scope="module")
def room():
    return sg.room_scene(width=24, height=18, spp=4, freq=8)


def _render(device, scene, p, area=AREA, depth=DEPTH):
    gs = kr.GpuScene(device, scene.to_desc())
    r = kr.Render(gs, scene.camera_struct(), p)
    if area is not None:
        r.set_profiler(area, depth)
    r.run(0, p.spp)
    out = r.read()
    sl = r.read_slices() if area is not None else None
    r.free()
    gs.free()
    return out, sl


def test_slices_match_oracle(device, room):
    p = room.params(seed=9)
    _, (L, bounce, recs) = _render(device, room, p)
    x0, y0, x1, y1 = AREA
    assert L.shape[0] == (x1 - x0) * (y1 - y0) * p.spp
    orc = ol.OracleScene(room, "f64", traversal=1)
    same, close, total = 0, 0, 0
    for s in range(L.shape[0]):
        k, px = s % p.spp, s // p.spp
        x, y = x0 + px % (x1 - x0), y0 + px // (x1 - x0)
        rec, oL, ob = orc.profile_path(p, x, y, k, DEPTH)
        if ob != bounce[s]:
            continue  # an FP32 / FP64 difference took the path elsewhere
        same += 1
        g = np.concatenate([recs[s]["o"], recs[s]["d"], recs[s]["f"], recs[s]["L"], recs[s]["theta"][:, None]],
                           axis=1).astype(np.float64)
        for j in range(DEPTH):
            total += 1
            ok = (np.allclose(g[j, 0:6], rec[j, 0:6], rtol=1e-3, atol=2e-3)
                  and np.allclose(g[j, 6:12], rec[j, 6:12], rtol=2e-3, atol=2e-3)
                  and abs(g[j, 12] - rec[j, 12]) < 1e-2
                  and recs[s]["shader"][j] == int(rec[j, 13]))
            close += ok
        if np.all(np.isfinite(oL)):
            np.testing.assert_allclose(L[s], oL, rtol=5e-3, atol=5e-3)
    assert same >= 0.95 * L.shape[0], same
    assert close >= 0.97 * total, (close, total)
    # the loop counter: Russian roulette ends most paths of the closed room at
    # its first chance (counter 4); none outlives maxBounce
    assert bounce.max() <= p.max_bounce + 1 and (bounce == 4).mean() > 0.5


def test_area_restriction_keeps_the_other_paths_identical(device, room):
    p = room.params(seed=3, spp=1)
    (rgb, valid), (L, bounce, _) = _render(device, room, p)
    (rgb0, valid0), _ = _render(device, room, p, area=None)
    x0, y0, x1, y1 = AREA
    inside = np.zeros((p.height, p.width), bool)
    inside[y0:y1, x0:x1] = True
    inside = inside.ravel()
    # the area's paths are the unprofiled render's, bit for bit
    assert np.array_equal(rgb[inside], rgb0[inside]) and np.array_equal(valid[inside], valid0[inside])
    # the rest are not traced (zero radiance)
    assert not rgb[~inside].any()
    # with 1 spp a slice's radiance is its pixel's sum
    fin = np.all(np.isfinite(L), axis=1)
    assert np.array_equal(L[fin].astype(np.float64), rgb[inside][fin])


def test_debug_kernel_records_nothing(device, room):
    p = room.params(seed=3, spp=2)
    p.kernel = kr.KERNEL_NORMAL
    (rgb, _), (L, bounce, recs) = _render(device, room, p)
    assert not recs["L"].any() and not bounce.any()
    assert L.max() <= 1.0 + 1e-6


def test_rejected_combinations(device, room):
    gs = kr.GpuScene(device, room.to_desc())
    r = kr.Render(gs, room.camera_struct(), room.params())
    with pytest.raises(kr.KrError):
        r.set_profiler((10, 0, 5, 4), 8)  # x1 < x0
    pos = np.asarray(room.positions)
    nrc = kr.Nrc(device, pos.min(axis=0) - 0.01, pos.max(axis=0) + 0.01)
    r.set_profiler(AREA, 4)
    with pytest.raises(kr.KrError):
        r.attach_nrc(nrc)
    r.set_profiler(None)
    r.attach_nrc(nrc)
    with pytest.raises(kr.KrError):
        r.set_profiler(AREA, 4)
    r.attach_nrc(None)
    # an area beyond the image is clamped to it
    r.set_profiler((0, 0, 0xFFFFFFFF, 0xFFFFFFFF), 1)
    assert r.profiler_size() == (room.width * room.height * room.spp, 1)
    r.free()
    nrc.free()
    gs.free()
