"""Plastic material (plastic_material.cpp:30-56): a Lambert lobe and a GGX lobe with the
dielectric Fresnel term (fresnel_effect.cpp:5-33), sampled as a two-lobe collection
(scattering_function_collection.cpp:48-148).  The oracle is pinned by the textbook
Fresnel equations, the normalisation of the collection's pdf and the agreement of its
sampler with its pdf; the GPU is compared with the oracle image for image."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import orc_py
from rainbow_amd import abi, scenes


def _lib():
    L = orc_py.lib()
    L.orc_fresnel_dielectric.argtypes = [C.c_float, C.c_float, C.c_float]
    L.orc_fresnel_dielectric.restype = C.c_float
    M = C.POINTER(abi.rbw_material)
    L.orc_material_pdf.argtypes = [M, abi.f32p, abi.f32p]
    L.orc_material_pdf.restype = C.c_float
    L.orc_material_eval.argtypes = [M, abi.f32p, abi.f32p, abi.f32p]
    L.orc_material_sample.argtypes = [M, abi.f32p, C.c_float, C.c_float, abi.f32p, abi.f32p, abi.f32p]
    L.orc_material_sample.restype = C.c_int
    return L


def fresnel_textbook(cos_i, n1, n2):
    """unpolarised Fresnel reflectance (Hecht), float64"""
    if cos_i < 0:
        n1, n2, cos_i = n2, n1, -cos_i
    sin_t = n1 / n2 * math.sqrt(max(0.0, 1 - cos_i * cos_i))
    if sin_t >= 1:
        return 1.0
    cos_t = math.sqrt(1 - sin_t * sin_t)
    rs = (n1 * cos_i - n2 * cos_t) / (n1 * cos_i + n2 * cos_t)
    rp = (n2 * cos_i - n1 * cos_t) / (n2 * cos_i + n1 * cos_t)
    return 0.5 * (rs * rs + rp * rp)


def test_dielectric_fresnel():
    L = _lib()
    eta = 1.49
    assert abs(L.orc_fresnel_dielectric(1.0, 1.0, eta) - ((eta - 1) / (eta + 1)) ** 2) < 1e-6
    assert L.orc_fresnel_dielectric(0.0, 1.0, eta) == pytest.approx(1.0, abs=1e-6)
    # total internal reflection from inside beyond the critical angle
    crit = math.sqrt(1 - 1 / eta ** 2)
    assert L.orc_fresnel_dielectric(-0.5 * crit, 1.0, eta) == 1.0
    for c in np.linspace(-1, 1, 41):
        assert abs(L.orc_fresnel_dielectric(float(c), 1.0, eta) - fresnel_textbook(float(c), 1.0, eta)) < 2e-6, c


def _plastic(kd, ks, alpha, ior=1.49):
    m = abi.rbw_material()
    m.type = abi.RBW_MATERIAL_PLASTIC
    m.diffuse[:] = kd
    m.specular[:] = ks
    m.alpha_u = m.alpha_v = alpha
    m.ior = ior
    return m


def _dir(theta):
    return np.array([math.sin(theta), 0.0, math.cos(theta)], np.float32)


def _pdf_integral(L, m, wo, nc=400, nphi=100):
    """midpoint quadrature of pdf(wo, .) over the upper hemisphere in (cos theta, phi)"""
    tot = 0.0
    for ci in (np.arange(nc) + 0.5) / nc:
        s = math.sqrt(1 - ci * ci)
        for p in (np.arange(nphi) + 0.5) / nphi * 2 * math.pi:
            wi = np.array([s * math.cos(p), s * math.sin(p), ci], np.float32)
            tot += L.orc_material_pdf(C.byref(m), abi.ptr(wo, abi.f32p), abi.ptr(wi, abi.f32p))
    return tot * (1.0 / nc) * (2 * math.pi / nphi)


def test_collection_pdf_averages_the_lobes():
    """Lambert's pdf integrates to 1; the GGX lobe's to the share of reflections that stay
    above the horizon; the two-lobe collection's pdf is their average."""
    L = _lib()
    alpha = 0.35
    matte = abi.rbw_material(type=abi.RBW_MATERIAL_MATTE)
    matte.diffuse[:] = (0.5, 0.4, 0.3)
    metal = abi.rbw_material(type=abi.RBW_MATERIAL_METAL, alpha_u=alpha, alpha_v=alpha)
    metal.eta[:] = (1.5, 1.5, 1.5)
    plastic = _plastic((0.5, 0.4, 0.3), (0.6, 0.6, 0.6), alpha)
    for th in (0.2, 0.9, 1.3):
        wo = _dir(th)
        lam = _pdf_integral(L, matte, wo)
        ggx = _pdf_integral(L, metal, wo)
        both = _pdf_integral(L, plastic, wo)
        assert abs(lam - 1.0) < 2e-3, (th, lam)
        assert 0.5 < ggx <= 1.0 + 2e-3
        assert abs(both - 0.5 * (lam + ggx)) < 2e-3, (th, both, lam, ggx)


def test_sampler_agrees_with_pdf_and_eval():
    """E[f cos / pdf] over the collection's own samples equals the quadrature of f cos;
    a sampler that picks lobes, or averages their pdfs, unlike pdf() would bias it."""
    L = _lib()
    m = _plastic((0.3, 0.5, 0.2), (0.8, 0.8, 0.8), 0.2)
    rng = np.random.default_rng(3)
    wo = _dir(0.7)
    wi = np.zeros(3, np.float32)
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    est, n = np.zeros(3), 40000
    for u0, u1 in rng.random((n, 2)):
        if L.orc_material_sample(C.byref(m), abi.ptr(wo, abi.f32p), float(u0), float(u1), abi.ptr(wi, abi.f32p),
                                 abi.ptr(f, abi.f32p), abi.ptr(pdf, abi.f32p)):
            p = float(pdf[0])
            # the returned value and pdf are the collection's, for the sampled direction
            assert p == pytest.approx(L.orc_material_pdf(C.byref(m), abi.ptr(wo, abi.f32p), abi.ptr(wi, abi.f32p)),
                                      rel=1e-5)
            est += f * abs(wi[2]) / p
    est /= n
    # quadrature of f cos over the hemisphere
    quad, nc, nphi = np.zeros(3), 500, 200
    for ci in (np.arange(nc) + 0.5) / nc:
        s = math.sqrt(1 - ci * ci)
        for ph in (np.arange(nphi) + 0.5) / nphi * 2 * math.pi:
            w = np.array([s * math.cos(ph), s * math.sin(ph), ci], np.float32)
            L.orc_material_eval(C.byref(m), abi.ptr(wo, abi.f32p), abi.ptr(w, abi.f32p), abi.ptr(f, abi.f32p))
            quad += f * ci
    quad *= (1.0 / nc) * (2 * math.pi / nphi)
    assert np.allclose(est, quad, rtol=0.02), (est, quad)


def plastic_room(res=48):
    sc = scenes.Scene()
    scenes.add_cornell_box(sc, spheres=False)
    sc.add_sphere((-0.45, -0.7, -0.35), 0.3, sc.plastic((0.6, 0.1, 0.1), (0.5, 0.5, 0.5), 0.08))
    sc.add_sphere((0.45, -0.75, 0.25), 0.25, sc.plastic((0.1, 0.3, 0.7), (0.9, 0.9, 0.9), 0.3, ior=1.8))
    cam = scenes.perspective_camera(res, res, 40.0, eye=(0.0, 0.0, 3.5), target=(0.0, 0.0, 0.0))
    return sc, cam


@pytest.mark.gpu
def test_gpu_plastic_matches_oracle(tracer):
    from rainbow_amd.tracer import render_params
    sc, cam = plastic_room()
    tracer.upload(sc)
    g = tracer.render(cam, render_params(spp=16, max_depth=5, seed=2))["image"]
    o = orc_py.OracleScene(sc).render(cam, orc_py.render_params(spp=16, max_depth=5, seed=2),
                                      sampler=orc_py.ORC_SAMPLER_PER_SAMPLE)["image"]
    bad = (np.abs(g - o) > 1e-4 * np.maximum(np.abs(g), np.abs(o)) + 1e-6).any(-1)
    assert bad.mean() < 1e-3, bad.sum()
