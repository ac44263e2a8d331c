"""The closed-form pins of tests/test_oracle_pins.py run on the HIP path (through the
C ABI): sphere intersection (rainbow-cpu/shapes/sphere.cpp:17-38) against exact
distances, Russian roulette (integrators/path_integrator.cpp:35-43) in furnaces
deep enough for it to fire, on the persistent-slot schedule, the pass schedule
and the guided estimator, and the thin-lens camera (cameras/perspective_camera.cpp:
78-99) against the oracle with lens_radius > 0.
"""
import ctypes as C

import numpy as np
import pytest

from rainbow_amd import abi, scenes
from rainbow_amd.tracer import guiding_opts, render_params
from test_oracle_pins import (RR_CASES, check_furnace, check_sphere_hits, closed_box, sphere_ray_sets,
                              sphere_scene)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("width", [2, 4, 8])
def test_gpu_sphere_closed_forms(tracer, width):
    tracer.upload(sphere_scene())
    for name, o, d, want in sphere_ray_sets():
        hits = tracer.trace(o, d, np.full(len(o), np.inf, np.float32), bvh_width=width)
        check_sphere_hits(name, o, d, want, hits)


@pytest.mark.parametrize("albedo,threshold,depth", RR_CASES)
def test_gpu_furnace_with_russian_roulette(tracer, albedo, threshold, depth):
    """persistent slots and passes of samples (the instrumented render, count_visits, runs
    there)"""
    sc, cam = closed_box(albedo, 1.0)
    tracer.upload(sc)
    for counted in (0, 1):
        r = tracer.render(cam, render_params(spp=256, max_depth=depth, rr_threshold=threshold, seed=7,
                                             count_visits=counted, schedule=0 if counted else 1))
        st = r["stats"]
        print("schedule", "passes" if counted else "persistent", st["extension_rays"], st["samples"])
        check_furnace(r["image"], st["extension_rays"], st["samples"], albedo, threshold, depth)


@pytest.mark.parametrize("guide_depth", [0, 1])
@pytest.mark.parametrize("albedo,threshold,depth", RR_CASES[:2])
def test_gpu_guided_furnace_with_russian_roulette(tracer, albedo, threshold, depth, guide_depth):
    """the guided vertex applies roulette to its own throughput (the mixture's f cos / p_G is
    not the albedo, so the ray count has no closed form): the expectation still holds"""
    sc, cam = closed_box(albedo, 1.0)
    tracer.upload(sc)
    tracer.guide_reset(5)
    go = guiding_opts(train=True, steps_per_pass=2, batch=1 << 14, guide_depth=guide_depth)
    r = tracer.render(cam, render_params(spp=256, max_depth=depth, rr_threshold=threshold, seed=8,
                                         guiding=C.pointer(go)))
    st = r["stats"]
    check_furnace(r["image"], st["extension_rays"], st["samples"], albedo, threshold, depth, check_rays=False)


def _pixel_mismatch(a, b):
    return (np.abs(a - b) > 1e-4 * np.maximum(np.abs(a), np.abs(b)) + 1e-7).any(axis=-1)


@pytest.mark.parametrize("size", [(64, 64), (80, 48)])
def test_gpu_thin_lens_camera_matches_the_oracle(tracer, size):
    """Cornell box through a thin lens focused on the spheres (focus plane in front of the
    lens: negative focus_distance, see test_oracle_thin_lens_rays_meet_at_the_focus_point),
    per-sample pcg32: radiance per pixel and every primary/shadow/MIS hit id"""
    from oracle import orc_py
    W, H = size
    sc, _ = scenes.cornell_box()
    cam = scenes.perspective_camera(W, H, 40.0, eye=(0.0, 0.0, 3.5), target=(0.0, -0.3, 0.0), lens_radius=0.15,
                                    focus=-3.6)
    tracer.upload(sc)
    p = orc_py.render_params(spp=16, max_depth=5, seed=13)
    c = orc_py.OracleScene(sc).render(cam, p, sampler=orc_py.ORC_SAMPLER_PER_SAMPLE, debug=True)
    g = tracer.render(cam, render_params(spp=16, max_depth=5, seed=13), debug=True)
    bad = _pixel_mismatch(g["image"], c["image"])
    gh, ch = g["debug_hits"], c["debug_hits"]
    traced = ch != abi.RBW_NONE
    ids = np.count_nonzero((gh != ch) & (traced | (gh != abi.RBW_NONE)))
    print(f"thin lens {W}x{H}: pixels off {bad.sum()} / {bad.size}, hit ids off {ids} / {traced.sum()}")
    assert bad.sum() <= max(1, bad.size // 1000)
    assert ids <= 1e-5 * traced.sum()
    # the persistent-slot schedule (no debug hits) gives the same image bit for bit
    p2 = tracer.render(cam, render_params(spp=16, max_depth=5, seed=13, schedule=1))["image"]
    assert np.array_equal(p2, g["image"])
