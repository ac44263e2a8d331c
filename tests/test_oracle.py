"""The CPU oracle pinned against the reference's own pcg32 build (golden
vectors), closed forms of the sampling/BSDF routines, brute force for the BVH,
and analytic integrals for the integrator (uniform furnace, polygon light)."""
import json
import math
import os

import numpy as np
import pytest

from oracle import orc_py
from rainbow_amd import abi, scenes

HERE = os.path.dirname(os.path.abspath(__file__))


def test_pcg32_matches_reference_golden():
    g = json.load(open(os.path.join(HERE, "golden", "pcg32.json")))
    for v in g["vectors"]:
        u = orc_py.pcg32_u32(v["seed"], g["n"])
        r = orc_py.pcg32_reals(v["seed"], g["n"])
        assert [int(x) for x in u] == v["u32"], v["seed"]
        assert [int(x) for x in r.view(np.uint32)] == v["real_bits"], v["seed"]


def test_pcg32_matches_reference_build():
    """Where the reference tree is present, compare with its pcg32 compiled in place."""
    if not os.path.exists(orc_py.REF_PCG):
        pytest.skip("oracle/_ref not built (no reference tree)")
    import ctypes as C
    lib = C.CDLL(orc_py.REF_PCG)
    lib.ref_pcg32_reals.argtypes = [C.c_ulonglong, C.c_uint, C.POINTER(C.c_float)]
    for seed in (3, 99, 123456789, 2**40 + 17):
        ref = np.zeros(1000, np.float32)
        lib.ref_pcg32_reals(seed, 1000, ref.ctypes.data_as(C.POINTER(C.c_float)))
        assert np.array_equal(ref.view(np.uint32), orc_py.pcg32_reals(seed, 1000).view(np.uint32))


def test_pcg32_uniform_real_range():
    r = orc_py.pcg32_reals(5, 200000)
    assert r.min() >= 0.0 and r.max() < 1.0
    assert abs(r.mean() - 0.5) < 0.005


def _call2(fn, u0, u1, n):
    out = np.zeros(n, np.float32)
    fn(np.float32(u0), np.float32(u1), abi.ptr(out, abi.f32p))
    return out


def test_concentric_disk_is_uniform_on_disk():
    L = orc_py.lib()
    rng = np.random.default_rng(0)
    pts = np.array([_call2(L.orc_concentric_sample_disk, a, b, 2) for a, b in rng.random((20000, 2))])
    r2 = (pts ** 2).sum(1)
    assert r2.max() <= 1.0 + 1e-6
    assert abs(r2.mean() - 0.5) < 0.01      # uniform disk: E[r^2] = 1/2
    assert abs(pts[:, 0].mean()) < 0.01 and abs(pts[:, 1].mean()) < 0.01
    # quadrant balance
    q = (pts[:, 0] > 0).astype(int) * 2 + (pts[:, 1] > 0)
    assert np.all(np.abs(np.bincount(q, minlength=4) / len(q) - 0.25) < 0.015)


def test_cosine_hemisphere_moments():
    L = orc_py.lib()
    rng = np.random.default_rng(1)
    w = np.array([_call2(L.orc_cosine_sample_hemisphere, a, b, 3) for a, b in rng.random((20000, 2))])
    assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-5)
    assert w[:, 2].min() >= 0
    assert abs(w[:, 2].mean() - 2.0 / 3.0) < 0.01   # E[cos] under pdf cos/pi


def test_uniform_triangle_barycentrics():
    L = orc_py.lib()
    rng = np.random.default_rng(2)
    b = np.array([_call2(L.orc_uniform_sample_triangle, a, c, 2) for a, c in rng.random((20000, 2))])
    assert b.min() >= 0 and (b.sum(1) <= 1 + 1e-6).all()
    assert np.allclose(b.mean(0), [1 / 3, 1 / 3], atol=0.01)


def _sph_grid(n_theta=400, n_phi=800):
    th = (np.arange(n_theta) + 0.5) / n_theta * (math.pi / 2)
    ph = (np.arange(n_phi) + 0.5) / n_phi * (2 * math.pi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    w = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dA = (np.sin(T) * (math.pi / 2 / n_theta) * (2 * math.pi / n_phi)).reshape(-1)
    return w.astype(np.float32), dA


@pytest.mark.parametrize("ax,ay", [(0.1, 0.1), (0.3, 0.3), (0.5, 0.2)])
def test_ggx_normalisation(ax, ay):
    """integral of D(h) cos(theta_h) over the hemisphere is 1"""
    L = orc_py.lib()
    w, dA = _sph_grid(600 if ax < 0.2 else 300, 600)
    D = np.array([L.orc_ggx_D(ax, ay, abi.ptr(np.ascontiguousarray(x), abi.f32p)) for x in w[::1]])
    integral = float((D * w[:, 2] * dA).sum())
    assert abs(integral - 1.0) < 0.01, integral


@pytest.mark.parametrize("alpha", [0.1, 0.4])
def test_ggx_sampling_pdf_consistent(alpha):
    """E[f cos / pdf] over visible-normal samples equals the quadrature of f cos (F = 1),
    and the pdf the sampler returns equals the pdf routine's value."""
    L = orc_py.lib()
    wo = np.array([0.3, -0.2, 0.0], np.float32)
    wo[2] = math.sqrt(1 - (wo[:2] ** 2).sum())
    rng = np.random.default_rng(3)
    est = []
    for u0, u1 in rng.random((20000, 2)):
        wi = np.zeros(3, np.float32)
        f = np.zeros(3, np.float32)
        pdf = np.zeros(1, np.float32)
        L.orc_ggx_sample_f(alpha, alpha, None, None, abi.ptr(wo, abi.f32p), float(u0), float(u1),
                           abi.ptr(wi, abi.f32p), abi.ptr(f, abi.f32p), abi.ptr(pdf, abi.f32p))
        if pdf[0] > 0:
            p2 = L.orc_ggx_pdf(alpha, alpha, abi.ptr(wo, abi.f32p), abi.ptr(wi, abi.f32p))
            assert abs(p2 - pdf[0]) <= 1e-3 * pdf[0]
            est.append(f[0] * abs(wi[2]) / pdf[0])
        else:
            est.append(0.0)
    mc = float(np.mean(est))
    w, dA = _sph_grid(500, 500)
    fq = np.zeros(len(w))
    for i in range(len(w)):
        fo = np.zeros(3, np.float32)
        L.orc_ggx_eval(alpha, alpha, None, None, abi.ptr(wo, abi.f32p), abi.ptr(np.ascontiguousarray(w[i]), abi.f32p),
                       abi.ptr(fo, abi.f32p))
        fq[i] = fo[0]
    quad = float((fq * w[:, 2] * dA).sum())
    assert abs(mc - quad) < 0.02 * quad, (mc, quad)
    assert quad <= 1.0 + 1e-3


def test_fresnel_conductor_normal_incidence():
    L = orc_py.lib()
    eta = np.array(scenes.GOLD_ETA, np.float32)
    k = np.array(scenes.GOLD_K, np.float32)
    out = np.zeros(3, np.float32)
    L.orc_fresnel_conductor(1.0, abi.ptr(eta, abi.f32p), abi.ptr(k, abi.f32p), abi.ptr(out, abi.f32p))
    e, kk = eta.astype(float), k.astype(float)
    closed = ((e - 1) ** 2 + kk ** 2) / ((e + 1) ** 2 + kk ** 2)
    assert np.allclose(out, closed, rtol=1e-5)
    L.orc_fresnel_conductor(0.0, abi.ptr(eta, abi.f32p), abi.ptr(k, abi.f32p), abi.ptr(out, abi.f32p))
    assert np.allclose(out, 1.0, atol=1e-5)   # grazing incidence reflects everything


def _random_rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.99, 0.99, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    return o, d


@pytest.mark.parametrize("visible", [False, True])
def test_oracle_bvh_equals_linear_scan(visible):
    sc = scenes.Scene()
    scenes.add_cornell_box(sc)
    P, F, N = scenes.fbm_displaced_icosphere(level=3, radius=0.4)
    sc.add_mesh(P, F, sc.matte((0.5, 0.5, 0.5)), normals=N)
    sc.add_mesh(P * 0.5 + np.float32(0.3), F, None)  # an invisible entity (no material)
    os_ = orc_py.OracleScene(sc)
    o, d = _random_rays(20000, 11)
    t = np.full(len(o), np.inf, np.float32)
    a = os_.trace(o, d, t, visible_only=visible)
    b = os_.trace(o, d, t, visible_only=visible, linear=True)
    assert np.count_nonzero(a["prim"] != b["prim"]) <= 2


def closed_box(albedo, Le):
    """six inward-facing emissive Lambert walls enclosing [-1,1]^3"""
    sc = scenes.Scene()
    m = sc.matte((albedo,) * 3)
    quads = [((-1, -1, 1), (1, -1, 1), (1, -1, -1), (-1, -1, -1)), ((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1)),
             ((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1)), ((-1, -1, 1), (-1, -1, -1), (-1, 1, -1), (-1, 1, 1)),
             ((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1)), ((1, -1, 1), (-1, -1, 1), (-1, 1, 1), (1, 1, 1))]
    for q in quads:
        P, I, N = scenes._quad(*q)
        sc.add_mesh(P, I, m, normals=N, radiance=(Le,) * 3)
    cam = scenes.perspective_camera(32, 32, 60.0, eye=(0.1, 0.2, 0.3), target=(0.5, -0.3, -1.0))
    return sc, cam


FURNACE_DEPTH = 3  # Russian roulette starts at bounce 4: none at this depth


def furnace_expected(albedo, Le, depth):
    # emission at the first vertex plus next-event estimation at vertices 0..depth-1
    return Le * sum(albedo ** k for k in range(depth + 1))


def test_oracle_uniform_furnace():
    sc, cam = closed_box(0.5, 1.0)
    os_ = orc_py.OracleScene(sc)
    r = os_.render(cam, orc_py.render_params(spp=64, max_depth=FURNACE_DEPTH, seed=4), sampler=orc_py.ORC_SAMPLER_PER_SAMPLE)
    mean = float(r["image"].mean())
    expect = furnace_expected(0.5, 1.0, FURNACE_DEPTH)
    assert abs(mean - expect) < 0.01 * expect, (mean, expect)


def polygon_irradiance(x, n, verts):
    """Lambert's formula: irradiance at x (normal n) from a unit-radiance polygon."""
    E = 0.0
    v = [np.asarray(p, float) - x for p in verts]
    for i in range(len(v)):
        a, b = v[i], v[(i + 1) % len(v)]
        c = np.cross(a, b)
        gamma = math.acos(np.clip(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b), -1, 1))
        E += gamma * np.dot(n, c / np.linalg.norm(c))
    return abs(E) / 2.0


def direct_light_scene():
    sc = scenes.Scene()
    P, I, N = scenes._quad((-5, 0, 5), (5, 0, 5), (5, 0, -5), (-5, 0, -5))
    sc.add_mesh(P, I, sc.matte((0.5, 0.5, 0.5)), normals=N)
    h = 0.25
    verts = [(-h, 1, -h), (h, 1, -h), (h, 1, h), (-h, 1, h)]
    P, I, N = scenes._quad(*verts)
    sc.add_mesh(P, I, sc.matte((0, 0, 0)), normals=N, radiance=(10.0, 10.0, 10.0))
    cam = scenes.perspective_camera(16, 16, 0.5, eye=(0, 0.3, 0), target=(0, 0, 0), up=(0, 0, -1))
    expect = 0.5 / math.pi * 10.0 * polygon_irradiance(np.zeros(3), np.array([0, 1.0, 0]), verts)
    return sc, cam, expect


def test_oracle_direct_lighting_polygon():
    sc, cam, expect = direct_light_scene()
    os_ = orc_py.OracleScene(sc)
    r = os_.render(cam, orc_py.render_params(spp=256, max_depth=1, seed=2), sampler=orc_py.ORC_SAMPLER_PER_SAMPLE)
    mean = float(r["image"][..., 0].mean())
    assert abs(mean - expect) < 0.01 * expect, (mean, expect)


def test_oracle_tile_stream_is_the_reference_seeding():
    """Tile mode: the first draws of the first sample of tile t are pcg32(t)'s (sampler_integrator.cpp:94)."""
    sc, cam = scenes.cornell_box()
    os_ = orc_py.OracleScene(sc)
    r = os_.render(cam, orc_py.render_params(spp=2, max_depth=5), sampler=orc_py.ORC_SAMPLER_TILE, record=True)
    offs, vals = r["replay_offsets"], r["replay_values"]
    W = cam.width
    for tile, (tx, ty) in enumerate([(0, 0), (1, 0), (0, 1), (5, 3)]):
        t_index = ty * (W // 16) + tx
        px = (ty * 16) * W + tx * 16
        sid = px * 2
        first = vals[offs[sid]: offs[sid] + 4]
        assert np.array_equal(first, orc_py.pcg32_reals(t_index, 4))
