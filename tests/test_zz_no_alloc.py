"""rbw_trace_rays_static, the entry point that allocates no device memory (its rays and
hit records live in a fixed per-device buffer), against rbw_trace_rays on the same rays:
hit records equal bit for bit, on two calls in a row (the second smaller, over the
buffer the first one left behind), at both BVH widths and both query kinds; sizes above
the buffer's limit are rejected."""
import numpy as np
import pytest

from rainbow_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.99, 0.99, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    t = np.where(rng.random(n) < 0.3, rng.uniform(0.05, 2.0, n), np.inf).astype(np.float32)
    return o, d, t


@pytest.fixture(scope="module")
def icosphere_scene(tracer):
    sc, _ = scenes.displaced_icosphere_box(level=6, width=64, height=64)
    tracer.upload(sc)
    return sc


@pytest.mark.parametrize("width", [2, 4, 8])
@pytest.mark.parametrize("visible", [False, True])
def test_static_trace_matches_trace(tracer, icosphere_scene, width, visible):
    for n, seed in ((abi.RBW_STATIC_MAX_RAYS, 1), (37_501, 2)):
        o, d, t = _rays(n, seed + 10 * width + visible)
        want = tracer.trace(o, d, t, visible_only=visible, bvh_width=width)
        got = tracer.trace(o, d, t, visible_only=visible, bvh_width=width, static=True)
        print(f"n {n}: hits {np.count_nonzero(got['prim'] != abi.RBW_MISS)}")
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def test_static_trace_rejects_more_rays_than_its_buffer(tracer, icosphere_scene):
    from rainbow_amd.tracer import RbwError
    o, d, t = _rays(abi.RBW_STATIC_MAX_RAYS + 1, 3)
    with pytest.raises(RbwError) as e:
        tracer.trace(o, d, t, static=True)
    assert e.value.code == abi.RBW_ERR_INVALID_ARGUMENT
    assert "RBW_STATIC_MAX_RAYS" in str(e.value)
    # an empty call is a no-op
    assert len(tracer.trace(o[:0], d[:0], t[:0], static=True)) == 0
