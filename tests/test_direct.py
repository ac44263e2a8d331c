"""The direct-lighting integrator (direct_integrator.cpp:20-128): the oracle restatement
pinned by Lambert's polygon-irradiance formula for several light/BSDF sample counts and
by its agreement with the path integrator at depth 1; the GPU compared with the oracle."""
import numpy as np
import pytest

from oracle import orc_py
from rainbow_amd import abi, scenes
from test_oracle import direct_light_scene


def _params(spp, ne, nb, seed=0, **kw):
    return orc_py.render_params(spp=spp, max_depth=5, seed=seed, integrator=abi.RBW_INTEGRATOR_DIRECT,
                                emitter_samples=ne, bsdf_samples=nb, **kw)


@pytest.mark.parametrize("ne,nb", [(1, 1), (2, 3), (4, 1), (1, 4)])
def test_oracle_direct_matches_polygon_irradiance(ne, nb):
    sc, cam, expect = direct_light_scene()
    o = orc_py.OracleScene(sc)
    img = o.render(cam, _params(256, ne, nb, seed=2), sampler=orc_py.ORC_SAMPLER_PER_SAMPLE)["image"]
    mean = float(img[..., 0].mean())
    assert abs(mean - expect) < 0.01 * expect, (ne, nb, mean, expect)


def test_oracle_direct_one_one_is_the_path_integrator_at_depth_one():
    """With one sample of each kind the weights are the path integrator's (the power
    heuristic is scale-free and 0.5 scales exactly) and both draw the camera, light and
    BSDF samples in the same order, so their images are the same bit for bit; at the
    light quad (a black matte: no lobes) the direct integrator passes through to the
    ceiling behind it, which sees none of the one-sided light."""
    sc, cam = scenes.cornell_box(32, 32)
    o = orc_py.OracleScene(sc)
    d = o.render(cam, _params(16, 1, 1, seed=3), sampler=orc_py.ORC_SAMPLER_PER_SAMPLE)["image"]
    p = o.render(cam, orc_py.render_params(spp=16, max_depth=1, seed=3), sampler=orc_py.ORC_SAMPLER_PER_SAMPLE)["image"]
    assert np.array_equal(d, p)


@pytest.mark.gpu
@pytest.mark.parametrize("ne,nb", [(1, 1), (2, 3)])
def test_gpu_direct_matches_oracle(tracer, ne, nb):
    from rainbow_amd.tracer import render_params
    sc, cam = scenes.cornell_box(48, 48)
    tracer.upload(sc)
    gp = render_params(spp=16, max_depth=5, seed=4, integrator=abi.RBW_INTEGRATOR_DIRECT, emitter_samples=ne,
                       bsdf_samples=nb)
    r = tracer.render(cam, gp)
    g = r["image"]
    o = orc_py.OracleScene(sc).render(cam, _params(16, ne, nb, seed=4), sampler=orc_py.ORC_SAMPLER_PER_SAMPLE)
    bad = (np.abs(g - o["image"]) > 1e-4 * np.maximum(np.abs(g), np.abs(o["image"])) + 1e-6).any(-1)
    assert bad.mean() < 1e-3, bad.sum()
    assert r["stats"]["shadow_rays"] == o["stats"]["shadow_rays"]
    assert r["stats"]["mis_rays"] == o["stats"]["mis_rays"]


@pytest.mark.gpu
def test_gpu_direct_rejects_guiding(tracer):
    import ctypes as C
    from rainbow_amd.tracer import RbwError, guiding_opts, render_params
    sc, cam = scenes.cornell_box(8, 8)
    tracer.upload(sc)
    go = guiding_opts()
    with pytest.raises(RbwError):
        tracer.render(cam, render_params(spp=4, integrator=abi.RBW_INTEGRATOR_DIRECT, guiding=C.pointer(go)))
