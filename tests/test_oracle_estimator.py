"""Independent pins of the oracle's multi-bounce estimator.

RadianceKernel::execute (src/core/kernel/radiance.cu:26-71) carries three
terms that no single-vertex check sees: the throughput update
beta *= f * |wi.z| (radiance.cu:56-58), the Russian roulette
q = max(.05, 1 - max beta), beta /= 1 - q after bounce 3 (radiance.cu:62-68),
and the BSDF sample weights it multiplies in (lambert.hpp:26-30,
glossy.hpp:24-27). Each test below fixes one of them from mathematics the
oracle does not contain (a closed-form integral, a quadrature of a textbook
density, unbiasedness of the roulette), so that a misreading shared by the
oracle and the GPU kernels would fail here:

- the Lambert sample weight: E[f |wi.z|] = color R * 2/3 over the
  cosine-distributed directions (int cos * cos/pi dw = 2/3);
- the Glossy same-hemisphere weight: the share of zero-weight samples equals
  the quadrature of P(wo.z wi.z <= 0) under the Beckmann density of normals
  D(h) cos(theta_h) in solid angle;
- the throughput update: with a maxBounce 1 path on a Lambert floor, the
  second vertex's emission adds color E cos(theta) per path whose bounce ray
  hits the light, so its mean is (color E / pi) * int_light cos^2(theta) dw,
  a 2-D quadrature over the light quad (a dropped cos doubles it here);
- Russian roulette: a ray through a stack of Transparent planes keeps a
  known throughput color^k; the share of paths that reach the emitter behind
  bounce 4 and the radiance they carry follow the roulette rule, and the mean
  equals color^6 E for any survival rule (unbiasedness).
"""
import ctypes as C
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import oracle_lib as ol  # noqa: E402
from koishi_amd import kr, scenegen as sg  # noqa: E402

D = C.c_double


def f32(x):
    """The FP32 value a shader parameter holds in kr_shader_desc."""
    return np.asarray(x, np.float32).astype(np.float64)


def _shader(cfg):
    return sg.shader_desc_from_config(cfg)


def _sample_grid(sh, wo, n):
    """oracle_shader_sample over the midpoints of an n x n grid of (u0, u1)."""
    f_s = ol.fn("oracle_shader_sample")
    wo3 = ol.d3(wo)
    wi, f = ol.d3(), ol.d3()
    WI = np.empty((n * n, 3))
    F = np.empty((n * n, 3))
    k = 0
    for a in range(n):
        u0 = (a + 0.5) / n
        for b in range(n):
            f_s(C.byref(sh), wo3, D(u0), D((b + 0.5) / n), D(0.5), D(0.0), None, wi, f)
            WI[k] = (wi[0], wi[1], wi[2])
            F[k] = (f[0], f[1], f[2])
            k += 1
    return WI, F


def test_lambert_sample_weight_mean():
    """lambert.hpp:26-30: f = color R Cosine::f(wi) / pdf with the cosine
    sampler's pdf = 1/pi. The integrator multiplies by |wi.z|
    (radiance.cu:57), so E[f |wi.z|] = color R E[cos] = color R * 2/3 for
    cosine-distributed wi, and f itself is the constant color R."""
    color, R = [0.6, 0.5, 0.3], [0.9, 0.8, 0.7]
    sh = _shader(["Lambert", {"color": color, "R": R}])
    WI, F = _sample_grid(sh, (0.3, -0.2, 0.9327379), 256)
    cr = f32(color) * f32(R)
    assert np.allclose(F, cr[None, :], rtol=1e-12, atol=0)
    mean = (F * np.abs(WI[:, 2:3])).mean(axis=0)
    assert np.allclose(mean, cr * 2.0 / 3.0, rtol=2e-4), (mean, cr * 2 / 3)


def _beckmann_zero_fraction(wo, alpha):
    """P(wo.z * wi.z <= 0) for wi = reflect(wo, h), h distributed with the
    Beckmann density of normals p(h) = D(h) cos(theta_h) in solid angle,
    D(h) = exp(-tan^2 theta / alpha^2) / (pi alpha^2 cos^4 theta): midpoint
    quadrature over (theta, phi). Returns the fraction and the integral of
    p(h) itself (which must be 1)."""
    wo = np.asarray(wo, np.float64)
    nt, nph = 3000, 1440
    th = (np.arange(nt) + 0.5) * (0.5 * math.pi / nt)
    ph = (np.arange(nph) + 0.5) * (2 * math.pi / nph)
    T, P = np.meshgrid(th, ph, indexing="ij")
    ct, st = np.cos(T), np.sin(T)
    a2 = alpha * alpha
    dens = np.exp(-(st / ct) ** 2 / a2) / (math.pi * a2 * ct ** 4) * ct * st  # p(h) * sin(theta) (area element)
    h = np.stack([st * np.cos(P), st * np.sin(P), ct], axis=-1)
    hw = (h * wo).sum(-1)
    wiz = 2 * hw * h[..., 2] - wo[2]
    dA = (0.5 * math.pi / nt) * (2 * math.pi / nph)
    total = dens.sum() * dA
    zero = (dens * (wo[2] * wiz <= 0)).sum() * dA
    return zero / total, total


@pytest.mark.parametrize("woz,alpha", [(0.25, 0.8), (0.15, 0.5), (0.5, 0.6)])
def test_glossy_sample_weight_is_zero_below_the_surface(woz, alpha):
    """glossy.hpp:24-27: wi = reflect(wo, h) and f = color only when wo and
    wi lie on one side of the surface (H::isSame), else 0. The share of
    zero-weight samples on a stratified grid must equal the quadrature of
    P(wo.z wi.z <= 0) under the Beckmann density, and the mean weight
    color * (1 - P)."""
    wo = np.array([math.sqrt(1 - woz * woz), 0.0, woz])
    color = [0.7, 0.4, 0.2]
    sh = _shader(["GlossyBSDF", {"distribution": ["Beckmann", {"roughness": alpha}], "color": color}])
    WI, F = _sample_grid(sh, wo, 320)
    zero = np.all(F == 0, axis=1)
    assert np.all(F[~zero] == f32(color))
    same = wo[2] * WI[:, 2] > 0
    assert np.array_equal(~zero, same)
    expect, total = _beckmann_zero_fraction(wo, alpha)
    assert abs(total - 1.0) < 2e-3
    assert expect > 0.05  # the case is not trivial
    assert abs(zero.mean() - expect) < 6e-3, (zero.mean(), expect)
    assert np.allclose(F.mean(axis=0), f32(color) * (1 - expect), atol=6e-3)


def _floor_light_scene(E=25.0, color=0.7):
    b = sg.Builder()
    b.shaders["floor"] = ["Lambert", {"color": [color, color, color]}]
    b.shaders["light"] = ["Emission", {"color": [1.0, 1.0, 1.0], "strength": E}]
    v, n, t = sg.quad((0.0, 0.0, 0.0), 8.0, 8.0, 1.0)
    b.add_mesh("floor", "floor", v, n, t)
    # a vertical 1 x 1 light to the side (Y-up: the x = 1.5 plane, facing -x),
    # so that the directions reaching it are far from the floor's normal
    v, n, t = sg.rect((1.5, 0.6, -0.5), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0), (-1.0, 0.0, 0.0))
    b.add_mesh("light", "light", v, n, t)
    return b.finish((0.0, 2.0, 0.0), (0.0, 0.0, 0.0), 30.0, 1, 1, 1, 1, "floor_light"), v


def _light_integral(x, quad_v, power=2, n=400):
    """int over the light quad of cos^power(theta_x) cos(theta_l) / r^2 dA,
    with theta_x measured from the floor normal (+z) and theta_l from the
    quad's normal: the solid-angle integral of cos^power over the light seen
    from x."""
    p0 = sg.rot_yup_to_zup(quad_v[0])
    e1 = sg.rot_yup_to_zup(quad_v[1]) - p0
    e2 = sg.rot_yup_to_zup(quad_v[3]) - p0
    nl = np.cross(e1, e2)
    area = np.linalg.norm(nl)
    nl /= area
    s = (np.arange(n) + 0.5) / n
    A, B = np.meshgrid(s, s, indexing="ij")
    pts = p0[None, None, :] + A[..., None] * e1 + B[..., None] * e2
    w = pts - x
    r2 = (w * w).sum(-1)
    wh = w / np.sqrt(r2)[..., None]
    cx = wh[..., 2]
    cl = np.abs((wh * nl).sum(-1))
    return float((cx ** power * cl / r2).mean() * area)


def test_second_vertex_weight_matches_quadrature():
    """The throughput update beta *= f |wi.z| (radiance.cu:57-58). On a
    Lambert floor of albedo c beside a vertical quad light of emission E, a
    path of maxBounce 1 differs from the same path at maxBounce 0 (same
    counter-based random numbers for vertex 0) by vertex 1 alone. When the
    bounce ray hits the light, vertex 1 adds beta E with
    beta = f |wi.z| = c cos(theta); the ray's direction has the cosine
    density cos(theta)/pi, so E[L1 - L0] = (c E / pi) int_light cos^2 dw.
    (The light sample at vertex 1 lies in the light's own plane and adds
    nothing.) Dropping the |wi.z| factor would make it (c E / pi) int cos dw,
    about twice as much for this light."""
    E, c = 25.0, 0.7
    s, lv = _floor_light_scene(E, c)
    orc = ol.OracleScene(s, "f64")
    x = np.array([0.0, 0.0, 0.0])
    o = ol.d3((0.0, 0.2, 1.0))
    dvec = np.array([0.0, -0.2, -1.0])
    d = ol.d3(dvec / np.linalg.norm(dvec))
    f = ol.fn("oracle_path_radiance")
    L = ol.d3()
    n = 120000
    diff = np.empty(n)
    for k in range(n):
        f(orc.h, o, d, 0, k, 77, 1, L)
        a = L[0]
        f(orc.h, o, d, 0, k, 77, 0, L)
        diff[k] = a - L[0]
    orc.free()
    # the camera ray meets the floor at (0, 0, 0) (Z-up: the floor is z = 0)
    expect = c * E / math.pi * _light_integral(x, lv)
    wrong = c * E / math.pi * _light_integral(x, lv, power=1)  # with beta *= f alone
    se = diff.std(ddof=1) / math.sqrt(n)
    assert (diff > 0).mean() > 0.01
    assert abs(diff.mean() - expect) <= 4 * se, (diff.mean(), expect, se)
    assert abs(wrong - expect) > 10 * se


def _transparent_stack(color, E, planes=6):
    b = sg.Builder()
    b.shaders["glass"] = ["Transparent", {"color": [color, color, color]}]
    b.shaders["light"] = ["Emission", {"color": [1.0, 1.0, 1.0], "strength": E}]
    for k in range(1, planes + 1):
        v, n, t = sg.quad((0.0, -float(k), 0.0), 20.0, 20.0, 1.0)
        b.add_mesh(f"plane{k}", "glass", v, n, t)
    v, n, t = sg.quad((0.0, -float(planes + 1), 0.0), 20.0, 20.0, 1.0)
    b.add_mesh("light", "light", v, n, t)
    return b.finish((0.0, 0.5, 0.0), (0.0, -1.0, 0.0), 30.0, 1, 1, 1, 8, "stack")


@pytest.mark.parametrize("color", [0.999, 0.8])
def test_russian_roulette_survival_and_compensation(color):
    """Russian roulette (radiance.cu:62-68). A ray straight down through six
    Transparent planes (f = color, wi = -wo, transparent.hpp:20-23) keeps
    beta = color^(b+1) after vertex b; the emitter plane behind them is
    vertex 6 and adds beta E. Roulette runs from bounce 4 when max beta < 1:
    q = max(.05, 1 - max beta), and a survivor's beta is divided by 1 - q.
    - color 0.999: q4 = .05 (the floor), survivors carry
      beta = color^5 / .95 > 1, so bounce 5 has no roulette: 95 % reach the
      emitter, each with L = color^6 E / .95.
    - color 0.8: q4 = 1 - .8^5, survivors carry beta = 1, bounce 5 then
      has q5 = .2 and leaves beta = 1 again: .8^5 * .8 of the paths reach
      the emitter, each with L = E.
    Either way the mean is color^6 E: the roulette is unbiased."""
    E = 5.0
    s = _transparent_stack(color, E)
    orc = ol.OracleScene(s, "f64")
    f = ol.fn("oracle_path_radiance")
    o = ol.d3((0.0, 0.0, 0.0))
    d = ol.d3((0.0, 0.0, -1.0))
    L = ol.d3()
    n = 40000
    vals = np.empty(n)
    for k in range(n):
        f(orc.h, o, d, 0, k, 5, 8, L)
        assert L[0] == L[1] == L[2]
        vals[k] = L[0]
    orc.free()
    c = float(f32(color))
    q_floor = float(f32(0.05))  # the reference's literal is the float .05f
    if color > 0.99:
        survive = 1 - q_floor
        carried = c ** 6 * E / (1 - q_floor)
    else:
        q4 = 1 - c ** 5
        q5 = 1 - c  # beta = c ** 5 / (1 - q4) * c = c
        survive = (1 - q4) * (1 - q5)
        carried = E
    alive = vals > 0
    assert np.allclose(vals[alive], carried, rtol=1e-12, atol=0), (vals[alive][:3], carried)
    se = math.sqrt(survive * (1 - survive) / n)
    assert abs(alive.mean() - survive) <= 4 * se, (alive.mean(), survive)
    mean_se = vals.std(ddof=1) / math.sqrt(n)
    assert abs(vals.mean() - c ** 6 * E) <= 4 * mean_se + 1e-12
