"""The stratified sampler (samplers/detail/stratified_sampler.hpp): the oracle's
restatement pinned by the reference's bounded pcg32 draws (golden vectors from its own
build), the generator's draw order and the stratification invariants; the oracle's
stratified render replayed on the GPU."""
import json
import os

import numpy as np
import pytest

from oracle import orc_py
from rainbow_amd import abi, scenes

HERE = os.path.dirname(os.path.abspath(__file__))


def test_bounded_draws_match_reference_golden():
    g = json.load(open(os.path.join(HERE, "golden", "pcg32.json")))
    for v in g["vectors"]:
        assert orc_py.pcg32_range(v["seed"], g["range_lo"], g["range_hi"]).tolist() == v["range"], v["seed"]


@pytest.mark.parametrize("strata,dims", [((4, 4), (3, 5)), ((3, 5), (2, 2)), ((1, 1), (1, 1)), ((8, 2), (0, 4))])
def test_tables_are_stratified_permutations(strata, dims):
    n = strata[0] * strata[1]
    t1, t2, draws = orc_py.stratified_tables(11, strata, dims)
    for d in range(dims[0]):
        k = np.floor(t1[d] * n).astype(int)
        assert sorted(k.tolist()) == list(range(n))          # one value per stratum
    for d in range(dims[1]):
        kx = np.floor(t2[d, :, 0] * strata[0]).astype(int)
        ky = np.floor(t2[d, :, 1] * strata[1]).astype(int)
        assert sorted((ky * strata[0] + kx).tolist()) == list(range(n))
    # n jitters (2n in 2D) and n bounded shuffle draws per dimension; a bounded draw
    # rejects with probability below n / 2^32
    assert draws == dims[0] * 2 * n + dims[1] * 3 * n


def test_draw_order_follows_the_generator():
    """1D dimensions first, jitters in stratum order before the shuffle's draws, then the
    2D dimensions with x before y per stratum, strata row by row."""
    strata, n = (2, 3), 6
    t1, t2, _ = orc_py.stratified_tables(5, strata, (1, 1))
    r = orc_py.pcg32_reals(5, 4 * n)
    assert np.allclose(sorted(t1[0]), sorted((np.arange(n) + r[:n]) / n), rtol=0, atol=1e-7)
    j = r[2 * n:4 * n].reshape(n, 2)
    want = [((i % 2 + j[i, 0]) / 2, (i // 2 + j[i, 1]) / 3) for i in range(n)]
    got = sorted(map(tuple, t2[0].tolist()))
    assert np.allclose(got, sorted(want), atol=1e-7)
    # the shuffle's first swap: position 0 takes stratum uint32(0, n-1), drawn right after the jitters
    u = orc_py.pcg32_u32(5, 2 * n)
    assert u[n] >= 4  # uint32(0, 5) rejects only draws below (2^32 - 6) % 6 = 4
    k0 = int(u[n] % n)
    assert int(np.floor(t1[0][0] * n)) == k0


def test_skip_moves_the_stream():
    a1, a2, d = orc_py.stratified_tables(3, (4, 4), (2, 2))
    b1, b2, _ = orc_py.stratified_tables(3, (4, 4), (2, 2), skip=d)
    assert not np.array_equal(a1, b1)


def _cornell(res=32):
    sc, cam = scenes.cornell_box(width=res, height=res)
    return sc, cam


def test_oracle_stratified_render_is_unbiased_and_less_noisy():
    sc, cam = _cornell(24)
    o = orc_py.OracleScene(sc)
    p = orc_py.render_params(spp=16, max_depth=2)
    ref = o.render(cam, orc_py.render_params(spp=4096, max_depth=2))["image"]
    rnd = o.render(cam, p)["image"]
    st = o.render(cam, p, sampler=orc_py.ORC_SAMPLER_TILE_STRATIFIED, strata=(4, 4), dims=(8, 8))["image"]
    e_rnd, e_st = scenes.rel_mse(rnd, ref), scenes.rel_mse(st, ref)
    assert e_st < 0.8 * e_rnd, (e_st, e_rnd)
    st256 = o.render(cam, orc_py.render_params(spp=256, max_depth=2), sampler=orc_py.ORC_SAMPLER_TILE_STRATIFIED,
                     strata=(16, 16), dims=(8, 8))["image"]
    assert abs(st256.mean() - ref.mean()) < 0.015 * ref.mean()


def test_oracle_rejects_mismatched_strata():
    sc, cam = _cornell(8)
    o = orc_py.OracleScene(sc)
    with pytest.raises(RuntimeError):
        o.render(cam, orc_py.render_params(spp=16, max_depth=2), sampler=orc_py.ORC_SAMPLER_TILE_STRATIFIED,
                 strata=(3, 3), dims=(2, 2))


@pytest.mark.gpu
def test_gpu_replays_the_stratified_stream(tracer):
    """The oracle's stratified render (the reference's sampler, fallback past the tables
    included) recorded and replayed on the GPU gives the same image."""
    from rainbow_amd.tracer import render_params
    sc, cam = _cornell(48)
    o = orc_py.OracleScene(sc)
    p = orc_py.render_params(spp=16, max_depth=5)
    r = o.render(cam, p, sampler=orc_py.ORC_SAMPLER_TILE_STRATIFIED, strata=(4, 4), dims=(3, 4), record=True)
    tracer.upload(sc)
    offs, vals = r["replay_offsets"], r["replay_values"]
    gp = render_params(spp=16, max_depth=5, sampler=abi.RBW_SAMPLER_REPLAY, replay_offsets=abi.ptr(offs, abi.u64p),
                       replay_values=abi.ptr(vals, abi.f32p), replay_count=len(vals))
    g = tracer.render(cam, gp)["image"]
    bad = (np.abs(g - r["image"]) > 1e-4 * np.maximum(np.abs(g), np.abs(r["image"])) + 1e-7).any(-1)
    assert bad.mean() < 1e-3, bad.sum()


def _strat_params(spp=16, strata=(4, 4), dims=(3, 4), seed=0, **kw):
    from rainbow_amd.tracer import render_params
    return render_params(spp=spp, seed=seed, sampler=abi.RBW_SAMPLER_STRATIFIED, strata_x=strata[0],
                         strata_y=strata[1], dims_1d=dims[0], dims_2d=dims[1], **kw)


@pytest.mark.gpu
def test_gpu_stratified_draws(tracer):
    """Each stratified dimension puts one sample in each stratum, independently permuted
    per dimension and pixel; draws past the stratified dimensions are the per-sample stream."""
    from rainbow_amd.tracer import render_params
    p = _strat_params(spp=12, strata=(4, 3), dims=(3, 4), seed=9)
    d = tracer.sampler_draws(p, 1234, 5, 6)
    assert d.shape == (12, 5 + 12) and (d >= 0).all() and (d < 1).all()
    for k in range(3):
        assert sorted(np.floor(d[:, k] * 12).astype(int).tolist()) == list(range(12))
    for k in range(4):
        x, y = d[:, 5 + 2 * k], d[:, 6 + 2 * k]
        cells = np.floor(y * 3).astype(int) * 4 + np.floor(x * 4).astype(int)
        assert sorted(cells.tolist()) == list(range(12))
    # independent permutations across dimensions and pixels
    assert not np.array_equal(np.argsort(d[:, 0]), np.argsort(d[:, 1]))
    d2 = tracer.sampler_draws(p, 1235, 5, 6)
    assert not np.array_equal(np.argsort(d[:, 0]), np.argsort(d2[:, 0]))
    # beyond the stratified dimensions: the per-sample pcg32 stream, as in the plain sampler
    pcg = tracer.sampler_draws(render_params(spp=12, seed=9), 1234, 0, 6)
    assert np.array_equal(d[:, 3:5], pcg[:, :2])
    assert np.array_equal(d[:, 5 + 8:], pcg[:, 2:6])
    assert np.array_equal(tracer.sampler_draws(p, 1234, 5, 6), d)  # deterministic


@pytest.mark.gpu
def test_gpu_stratified_render_unbiased_and_less_noisy(tracer):
    from rainbow_amd.tracer import render_params
    sc, cam = _cornell(128)
    tracer.upload(sc)
    ref = tracer.render(cam, render_params(spp=16384, max_depth=2, seed=77))["image"]
    e_p, e_s, m_p, m_s = [], [], [], []
    for seed in range(6):
        ip = tracer.render(cam, render_params(spp=16, max_depth=2, seed=seed))["image"]
        st = tracer.render(cam, _strat_params(spp=16, strata=(4, 4), dims=(8, 12), seed=seed, max_depth=2))["image"]
        e_p.append(scenes.rel_mse(ip, ref))
        e_s.append(scenes.rel_mse(st, ref))
        m_p.append(ip.mean())
        m_s.append(st.mean())
    assert np.mean(e_s) < 0.8 * np.mean(e_p), (np.mean(e_s), np.mean(e_p))
    assert abs(np.mean(m_s) - ref.mean()) < 0.01 * ref.mean(), (np.mean(m_s), ref.mean())


@pytest.mark.gpu
def test_gpu_stratified_schedules_agree(tracer):
    """Persistent slots and passes carry the per-sample dimension counters alike."""
    sc, cam = _cornell(48)
    tracer.upload(sc)
    a = tracer.render(cam, _strat_params(spp=16, dims=(2, 3), seed=4, max_depth=5, schedule=1))["image"]
    b = tracer.render(cam, _strat_params(spp=16, dims=(2, 3), seed=4, max_depth=5, count_visits=1))["image"]
    assert np.array_equal(a, b)


@pytest.mark.gpu
def test_gpu_stratified_params_are_validated(tracer):
    from rainbow_amd.tracer import RbwError
    sc, cam = _cornell(8)
    tracer.upload(sc)
    with pytest.raises(RbwError) as e:
        tracer.render(cam, _strat_params(spp=16, strata=(3, 3)))
    assert e.value.code == abi.RBW_ERR_INVALID_ARGUMENT
