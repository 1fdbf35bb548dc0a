"""The rest of the reference's Radiance-path plugin surface on the GPU,
against the FP64 oracle on identical seeded paths: GlassBSDF,
RefractionBSDF (Beckmann and Cosine), Transparent, Blackhole, Mixed with
Constant and Fresnel facs (nested, with an Emission leaf), point lights, and
the Normal / Visibility / HalfVector kernels."""
import numpy as np
import pytest

import oracle_lib as ol
from koishi_amd import kr, scenegen as sg
from test_gpu_parity import _check_paths, _per_path

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def materials():
    return sg.materials_scene()


def test_radiance_per_path_materials_scene(device, materials):
    out = _per_path(device, materials, 128, 72, [0, 3, 9], seed=21)
    _check_paths(out)


@pytest.mark.parametrize("kernel", [kr.KERNEL_NORMAL, kr.KERNEL_VISIBILITY, kr.KERNEL_HALFVECTOR])
def test_debug_kernels(device, materials, kernel):
    gs = kr.GpuScene(device, materials.to_desc())
    orc = ol.OracleScene(materials, "f64", traversal=1)
    p = materials.params(seed=5, width=128, height=72)
    p.kernel = kernel
    r = kr.Render(gs, materials.camera_struct(), p)
    for k in (0, 6):
        r.reset()
        r.run(k, 1)
        g, gv = r.read()
        c, cv, _ = orc.render(p, k, 1)
        assert (gv == cv).mean() >= 0.999
        ok = (gv == 1) & (cv == 1)
        d = np.abs(g[ok] - c[ok]).max(axis=1)
        # Normal: the FP32 watertight barycentrics against the oracle's FP64
        # Moller-Trumbore ones move the interpolated normal by ~1e-5
        tol = {kr.KERNEL_NORMAL: 2e-4, kr.KERNEL_VISIBILITY: 0.0, kr.KERNEL_HALFVECTOR: 1e-3}[kernel]
        assert (d <= tol).mean() >= 0.999, (d > tol).mean()
        if kernel == kr.KERNEL_VISIBILITY:
            assert 0.05 < g[ok].mean() < 0.95  # both lit and shadowed pixels
    r.free()
    gs.free()


def test_point_lights_only(device):
    """No area light at all: a point light lights the scene on its own
    (li = color * strength, no falloff, src/core/light/pointLight.hpp:20-27)."""
    b = sg.Builder()
    b.shaders["floor"] = ["Lambert", {"color": [0.5, 0.5, 0.5]}]
    v, n, t = sg.quad((0.0, 0.0, 0.0), 8.0, 8.0, 1.0)
    b.add_mesh("floor", "floor", v, n, t)
    b.add_sphere((0.0, 0.7, 0.0), 0.7, "floor")
    b.add_point_light((1.0, 3.0, -1.0), [1.0, 0.8, 0.6], 4.0)
    s = b.finish((0.0, 2.0, -5.0), (0.0, 0.5, 0.0), 60.0, 96, 64, 4, 4, "points")
    out = _per_path(device, s, 96, 64, [0, 1], seed=3)
    _check_paths(out)
    assert np.concatenate([o[0] for o in out]).mean() > 0.01


def test_mixed_nesting_limit_is_enforced(device):
    s = sg.materials_scene()
    names, sh, n = s.shader_descs()
    # chain four Mixed levels by hand: 3 is the limit
    arr = (kr.ShaderDesc * 6)()
    for i in range(4):
        arr[i] = kr.ShaderDesc()
        arr[i].type = kr.SHADER_MIXED
        arr[i].fac_value = 0.5
        arr[i].child[0] = arr[i].child[1] = i + 1
    arr[4] = kr.ShaderDesc()
    arr[4].type = kr.SHADER_LAMBERT
    holder = s.to_desc()
    holder.desc.shaders = arr
    holder.desc.num_shaders = 5
    prims = (kr.PrimDesc * holder.desc.num_prims)()
    for i in range(holder.desc.num_prims):
        prims[i] = holder.desc.prims[i]
        prims[i].shader = 0
    holder.desc.prims = prims
    holder.keep += [arr, prims]
    with pytest.raises(kr.KrError) as e:
        kr.GpuScene(device, holder)
    assert e.value.code == 5  # KR_E_UNSUPPORTED
