"""The oracle's profiler slices (RadianceKernel::Slice, src/core/kernel/
radiance.hpp:17-55 recorded by radiance.cu:38-50, 74-78), pinned by geometry
that fixes theta, f and L independently of the recording code, by the path
radiance the oracle renders for the same path, and by the loop-counter
semantics of Russian roulette in a closed room."""
import math

import numpy as np
import pytest

from koishi_amd import scenegen as sg
from oracle_lib import OracleScene, lens_rays

GROUND = [0.5, 0.4, 0.3]
E = 20.0


def _plane_scene(light_at):
    """A Lambert ground (y = 0, Y-up) seen obliquely by a 1x1 camera, and one
    point light placed by light_at(p, wo) (Z-up world)."""
    eye, look = (0.0, 3.0, -3.0), (0.0, 0.0, 0.0)

    def build(light_zup):
        b = sg.Builder()
        b.shaders["ground"] = ["Lambert", {"color": GROUND}]
        v, n, t = sg.rect((-20.0, 0.0, -20.0), (40.0, 0.0, 0.0), (0.0, 0.0, 40.0), (0.0, 1.0, 0.0))
        b.add_mesh("ground", "ground", v, n, t)
        # point lights are given Y-up: undo R_x(+90) (x, y, z) -> (x, -z, y)
        lx, ly, lz = light_zup
        b.add_point_light((lx, lz, -ly), [1.0, 1.0, 1.0], E)
        return b.finish(eye, look, 60.0, 1, 1, 1, 0, "plane")

    probe = build((0.0, 0.0, 10.0))
    o, d = lens_rays(probe, probe.params())
    o, d = o[0], d[0] / np.linalg.norm(d[0])
    p = o + d * (-o[2] / d[2])  # the ground is z = 0 after the rotation
    wo = -d
    return build(light_at(p, wo)), p, wo


def _slice(scene, depth=4):
    osc = OracleScene(scene)
    rec, L, bounce = osc.profile_path(scene.params(), 0, 0, 0, depth)
    return osc, rec, L, bounce


def test_theta_zero_for_the_mirror_direction():
    # light along the mirror direction of wo: reflect(wi, n) = wo, theta = 0
    scene, p, wo = _plane_scene(lambda p, wo: p + 3.0 * np.array([-wo[0], -wo[1], wo[2]]))
    _, rec, L, bounce = _slice(scene)
    assert rec[0, 12] < 1e-5
    # Lambert eval f = color / pi (lambert.hpp:21-35; the scene holds the
    # color in FP32); li = E for a point light; L = f E |wi.z| with
    # wi.z = n . wi = wo.z
    f = np.float32(GROUND).astype(np.float64) / math.pi
    np.testing.assert_allclose(rec[0, 6:9], f, rtol=1e-12)
    # (the light position is FP32 in the scene: wi agrees with wo to ~1e-7)
    np.testing.assert_allclose(rec[0, 9:12], f * E * wo[2], rtol=1e-6)
    # the record's ray is the camera ray; the shader is the ground's (id 0)
    o, d = lens_rays(scene, scene.params())
    np.testing.assert_allclose(rec[0, 0:3], o[0], atol=1e-12)
    np.testing.assert_allclose(rec[0, 3:6], d[0], atol=1e-12)
    assert rec[0, 13] == 0
    # the BSDF-sampled ray leaves the open plane: one vertex, L unchanged
    assert bounce == 1
    np.testing.assert_allclose(L, rec[0, 9:12], rtol=1e-12)
    assert not rec[1:].any()


def test_theta_twice_the_incidence_for_a_light_behind_the_camera():
    # light along wo itself: reflect(wi, n) . wo = cos(2 theta_o)
    scene, p, wo = _plane_scene(lambda p, wo: p + 2.0 * wo)
    _, rec, L, bounce = _slice(scene)
    theta_o = math.acos(wo[2])
    assert rec[0, 12] == pytest.approx(2 * theta_o, abs=1e-5)
    assert rec[0, 12] > 1.0  # a sign error in reflect would give ~0 here


def test_slice_radiance_is_the_rendered_path_and_records_accumulate():
    scene = sg.room_scene(width=12, height=9, spp=2, freq=6)
    osc = OracleScene(scene)
    prm = scene.params()
    rgb, valid, _ = osc.render(prm)
    for x, y in ((3, 2), (7, 5), (10, 8)):
        acc = np.zeros(3)
        for k in range(prm.spp):
            rec, L, bounce = osc.profile_path(prm, x, y, k, depth=16)
            if np.all(np.isfinite(L)):
                acc += L
            n = min(bounce + 1, 16)
            Ls = rec[:n, 9:12]
            # the recorded L never decreases along the path (all terms >= 0) and
            # ends at or below the path radiance, which equals it when the last
            # recorded vertex is the last one of the loop
            assert np.all(np.diff(Ls, axis=0) >= -1e-12)
            assert np.all(Ls[-1] <= L + 1e-12)
            assert not rec[n:].any()
        np.testing.assert_allclose(acc, rgb[y * prm.width + x], rtol=1e-12, atol=1e-12)


def test_loop_counter_follows_russian_roulette():
    # closed room, maxBounce 8: no path escapes, so the loop ends by Russian
    # roulette (allowed from bounce 4 on, break before ++bounce: counter
    # 4..8) or after vertex 8 (counter 9)
    scene = sg.room_scene(width=8, height=6, spp=4, freq=6)
    osc = OracleScene(scene)
    prm = scene.params()
    counts = []
    for y in range(prm.height):
        for x in range(prm.width):
            for k in range(prm.spp):
                _, _, b = osc.profile_path(prm, x, y, k, depth=1)
                counts.append(b)
    counts = np.array(counts)
    assert counts.min() == 4 and counts.max() == 9
    assert (counts == 4).sum() > 0.2 * counts.size
