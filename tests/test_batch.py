"""Batched jobs (one wavefront for many renders of uneven size; no counterpart in
the reference).  The bar: every job's image is bit-identical to a single
rbw_render of the same scene, camera and parameters, whatever the path-pool
size, block size or graph capture."""
import numpy as np
import pytest

from rainbow_amd import scenes

SMALL = dict(tris=(1e3, 3e4), res=(24, 80), spp=(1, 12), depth=(3, 12))


def test_batch_workload_ranges():
    jobs = scenes.batch_workload(n_jobs=12, seed=2, **SMALL)
    for j in jobs:
        assert 0.8 * SMALL["tris"][0] <= j["triangles"] - 10 <= 1.2 * SMALL["tris"][1]
        assert SMALL["res"][0] <= j["camera"].width <= SMALL["res"][1]
        assert j["camera"].width == j["camera"].height
        assert SMALL["spp"][0] <= j["spp"] <= SMALL["spp"][1]
        assert SMALL["depth"][0] <= j["max_depth"] <= SMALL["depth"][1]
    # seeded: the same list again
    again = scenes.batch_workload(n_jobs=12, seed=2, **SMALL)
    assert [(j["triangles"], j["spp"], j["max_depth"]) for j in jobs] == \
           [(j["triangles"], j["spp"], j["max_depth"]) for j in again]


def test_grid_sphere_is_closed_and_outward():
    P, F, N = scenes.fbm_displaced_grid_sphere(5000, amplitude=0.0, center=(0.0, 0.0, 0.0))
    P = P.astype(np.float64)
    fn = np.cross(P[F[:, 1]] - P[F[:, 0]], P[F[:, 2]] - P[F[:, 0]])
    c = P[F].mean(axis=1)
    assert np.all((fn * c).sum(1) > 0)
    # closed 2-manifold: every edge shared by exactly two faces, Euler characteristic 2
    e = np.sort(np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]), axis=1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    assert np.all(counts == 2)
    assert len(P) - len(counts) + len(F) == 2


def _submit(tracer, jobs, ids, **opts):
    from rainbow_amd.tracer import batch_opts
    js = [{"scene": ids[i], "camera": j["camera"], "spp": j["spp"], "max_depth": j["max_depth"],
           "seed": j["seed"]} for i, j in enumerate(jobs)]
    return tracer.submit_jobs(js, batch_opts(**opts))


@pytest.fixture(scope="module")
def small_batch(tracer):
    jobs = scenes.batch_workload(n_jobs=6, seed=2, **SMALL)
    ids = [tracer.add_scene(j["scene"]) for j in jobs]
    return jobs, ids


@pytest.mark.gpu
def test_batch_matches_single_renders(tracer, small_batch):
    from rainbow_amd.tracer import render_params
    jobs, ids = small_batch
    imgs, jstats, bstats = _submit(tracer, jobs, ids)
    print("batch", bstats)
    assert bstats["samples"] == sum(j["camera"].width * j["camera"].height * j["spp"] for j in jobs)
    assert bstats["graph_launches"] > 0
    for i, j in enumerate(jobs):
        tracer.upload(j["scene"])
        single = tracer.render(j["camera"], render_params(spp=j["spp"], max_depth=j["max_depth"], seed=j["seed"]),
                               stats=False)["image"]
        assert np.isfinite(single).all()
        assert single.mean() > 0
        assert np.array_equal(imgs[i], single), (i, np.abs(imgs[i] - single).max())
        js = jstats[i]
        assert js["samples"] == j["camera"].width * j["camera"].height * j["spp"]
        assert 1 <= js["first_iteration"] <= js["last_iteration"] <= bstats["iterations"]


@pytest.mark.gpu
@pytest.mark.parametrize("opts", [dict(use_graph=False), dict(pool_size=1024, block_size=64),
                                  dict(pool_size=1 << 12, block_size=512), dict(pool_size=3001, block_size=128)])
def test_batch_independent_of_schedule(tracer, small_batch, opts):
    """the path-pool size, block size and graph capture change the schedule, not the images
    (3001 slots: the persistent extension walk's last slot range and the last group are
    partial)"""
    jobs, ids = small_batch
    ref, _, _ = _submit(tracer, jobs, ids)
    imgs, _, bstats = _submit(tracer, jobs, ids, **opts)
    print(opts, bstats)
    for a, b in zip(ref, imgs):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_batch_shared_scene_and_autotune(tracer, small_batch):
    """jobs may share a scene; the autotuner picks a candidate configuration"""
    from rainbow_amd.tracer import batch_opts
    jobs, ids = small_batch
    cam2 = scenes.perspective_camera(40, 30, 30.0, eye=(0.5, 0.2, 3.0), target=(0.0, -0.2, 0.0))
    js = [{"scene": ids[0], "camera": jobs[0]["camera"], "spp": 3, "max_depth": 5, "seed": 1},
          {"scene": ids[0], "camera": cam2, "spp": 2, "max_depth": 4, "seed": 2},
          {"scene": ids[1], "camera": jobs[1]["camera"], "spp": 1, "max_depth": 6, "seed": 3}]
    o, ms = tracer.autotune_jobs(js)
    assert o.block_size in (64, 128, 256, 512) and o.pool_size in (1 << 20, 1 << 21, 1 << 22)
    assert o.bvh_width in (2, 4, 8) and ms > 0
    imgs, _, st = tracer.submit_jobs(js, batch_opts(autotune=True))
    assert st["autotune_ms"] > 0
    ref, _, _ = tracer.submit_jobs(js, batch_opts())
    for a, b in zip(ref, imgs):
        assert np.array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("width", [4, 8])
def test_batch_wide_bvh_matches_single_renders(tracer, small_batch, width):
    """the four- and eight-wide BVHs in the merged scene (rebased wide nodes and leaf slots)
    against single renders with the same width"""
    from rainbow_amd.tracer import render_params
    jobs, ids = small_batch
    imgs, _, st = _submit(tracer, jobs, ids, bvh_width=width)
    assert st["bvh_width"] == width
    for i, j in enumerate(jobs):
        tracer.upload(j["scene"])
        single = tracer.render(j["camera"], render_params(spp=j["spp"], max_depth=j["max_depth"], seed=j["seed"],
                                                          bvh_width=width), stats=False)["image"]
        assert np.array_equal(imgs[i], single), i


@pytest.mark.gpu
def test_batch_rejects_bad_jobs(tracer, small_batch):
    from rainbow_amd.tracer import RbwError
    jobs, ids = small_batch
    with pytest.raises(RbwError):
        _submit(tracer, jobs[:1], [12345])
    bad = dict(jobs[0])
    bad["spp"] = 0
    with pytest.raises(RbwError):
        _submit(tracer, [bad], ids[:1])
    with pytest.raises(RbwError):
        _submit(tracer, jobs[:1], ids[:1], block_size=100)


@pytest.mark.gpu
def test_batch_64_jobs_match_single_renders(tracer):
    """The batched-jobs workload at its configured size: 64 seeded jobs (seed 2) of
    1e4-2e6 triangles (25.6 M in all), 128^2-1024^2, 4-256 spp, depth 3-12 in one
    submission (the autotuner's choice of schedule).  Every job's image equals its single
    rbw_render with the BVH width the batch ran (oracle-pinned by the parity tests) bit for
    bit, and every job's samples are accounted for."""
    import time
    from rainbow_amd.tracer import batch_opts, render_params
    t0 = time.perf_counter()
    jobs = scenes.batch_workload()
    ids = [tracer.add_scene(j["scene"]) for j in jobs]
    print(f"64 jobs built and pooled in {time.perf_counter() - t0:.1f} s", flush=True)
    js = [{"scene": ids[i], "camera": j["camera"], "spp": j["spp"], "max_depth": j["max_depth"], "seed": j["seed"]}
          for i, j in enumerate(jobs)]
    try:
        imgs, jstats, bst = tracer.submit_jobs(js, batch_opts(autotune=True))
        print("batch", bst, flush=True)
        assert bst["samples"] == sum(j["camera"].width * j["camera"].height * j["spp"] for j in jobs) == 1099077143
        for i, j in enumerate(jobs):
            tracer.upload(j["scene"])
            single = tracer.render(j["camera"], render_params(spp=j["spp"], max_depth=j["max_depth"], seed=j["seed"],
                                                              bvh_width=bst["bvh_width"]), stats=False)["image"]
            assert np.isfinite(single).all() and single.mean() > 0, i
            assert np.array_equal(imgs[i], single), (i, np.abs(imgs[i] - single).max())
            assert jstats[i]["samples"] == j["camera"].width * j["camera"].height * j["spp"]
    finally:
        for sid in ids:
            tracer.remove_scene(sid)
    print(f"64 jobs checked in {time.perf_counter() - t0:.1f} s", flush=True)


@pytest.mark.gpu
def test_persistent_extension_walk_matches_one_slot_per_lane(tracer):
    """A render whose pool gives every resident wave a slot range (1024x1024 x 4 spp: 4.2 M
    slots) runs the persistent four-wide extension walk and the persistent occlusion walk
    (shadow rays and the queued MIS rays that reach an emitter triangle) that refill idle
    lanes, and shades each 256-slot window grouped by material kind (the icosphere's
    triangles are half Lambert, half GGX); the same render with visit counting runs the pass
    schedule, one ray per lane, in slot order.  Each ray's walk and each vertex's shading are
    the same, so the images and the ray counts are equal bit for bit."""
    from rainbow_amd.tracer import render_params
    sc, cam = scenes.displaced_icosphere_box(level=6, width=1024, height=1024)
    tracer.upload(sc)
    a = tracer.render(cam, render_params(spp=4, max_depth=8, seed=21, schedule=1))
    b = tracer.render(cam, render_params(spp=4, max_depth=8, seed=21, count_visits=1))
    assert a["stats"]["bvh_width"] == 4 and b["stats"]["node_visits"] > 0
    assert np.isfinite(a["image"]).all() and a["image"].mean() > 0
    assert np.array_equal(a["image"], b["image"])
    for k in ("extension_rays", "shadow_rays", "mis_rays", "light_casts"):
        print(k, a["stats"][k], b["stats"][k])
        assert a["stats"][k] == b["stats"][k], k
