"""Path guiding: the float64 oracle pinned by finite differences, closed forms
(vMF normalisation, Adam's first step) and sampling histograms; the HIP
FP16-MFMA kernels (inference, gradients, Adam) checked against the oracle;
the guided estimator's unbiasedness checked on the GPU."""
import math

import numpy as np
import pytest

from oracle import guiding_ref as G
from rainbow_amd import abi, scenes


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 9))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = rng.random(n) + 0.1
    return x, d, c


def test_oracle_parameter_count():
    assert G.n_params() == 33128          # 216*64+64 + 4*(64*64+64) + 64*40+40
    assert G.to_padded(*G.init_params(0)).size == abi.RBW_GUIDE_PARAMS


def test_oracle_gradients_match_finite_differences():
    Ws, bs = G.init_params(1)
    x, d, c = _batch(48, 0)
    X = G.encode(x)
    o, acts, pre = G.forward(Ws, bs, X)
    _, dout = G.loss_and_grad_out(o, d, c)
    dWs, dbs = G.backward(Ws, bs, acts, pre, dout)
    eps = 1e-6
    rng = np.random.default_rng(3)
    for l in range(6):
        for _ in range(4):
            a = rng.integers(Ws[l].shape[0])
            b = rng.integers(G.N_FEAT if l == 0 else Ws[l].shape[1])
            Wp = [w.copy() for w in Ws]
            Wm = [w.copy() for w in Ws]
            Wp[l][a, b] += eps
            Wm[l][a, b] -= eps
            fd = (G.loss_fn(Wp, bs, X, d, c) - G.loss_fn(Wm, bs, X, d, c)) / (2 * eps)
            assert abs(fd - dWs[l][a, b]) <= 1e-6 + 1e-4 * abs(fd), (l, a, b, fd, dWs[l][a, b])
        j = rng.integers(len(bs[l]))
        bp = [v.copy() for v in bs]
        bm = [v.copy() for v in bs]
        bp[l][j] += eps
        bm[l][j] -= eps
        fd = (G.loss_fn(Ws, bp, X, d, c) - G.loss_fn(Ws, bm, X, d, c)) / (2 * eps)
        assert abs(fd - dbs[l][j]) <= 1e-6 + 1e-4 * abs(fd)


@pytest.mark.parametrize("kappa", [0.05, 1.0, 30.0, 400.0])
def test_vmf_normalised(kappa):
    n = 400
    th = (np.arange(n) + 0.5) / n * math.pi
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * math.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dA = (np.sin(T) * (math.pi / n) * (math.pi / n)).reshape(-1)
    mu = np.tile([0.0, 0.0, 1.0], (len(d), 1))
    pdf = G.vmf_pdf(d, mu, np.full(len(d), kappa))
    assert abs((pdf * dA).sum() - 1.0) < 2e-3


def test_vmf_sampler_matches_pdf():
    rng = np.random.default_rng(5)
    mu = np.array([0.3, -0.5, 0.8])
    mu /= np.linalg.norm(mu)
    kappa = 12.0
    cos = np.array([G.vmf_sample(mu, kappa, *rng.random(2)) @ mu for _ in range(40000)])
    # the cosine to the mean has density kappa e^{kappa (t-1)} / (1 - e^{-2 kappa}) on [-1, 1]
    edges = np.linspace(0.5, 1.0, 11)
    hist, _ = np.histogram(cos, edges)
    cdf = lambda t: (np.exp(kappa * (t - 1)) - np.exp(-2 * kappa)) / (1 - np.exp(-2 * kappa))
    expect = (cdf(edges[1:]) - cdf(edges[:-1])) * len(cos)
    assert np.all(np.abs(hist - expect) < 5 * np.sqrt(expect) + 5)


def test_adam_first_step_closed_form():
    p = [np.array([1.0, -2.0, 0.5])]
    g = [np.array([0.3, -1e-3, 4.0])]
    opt = G.Adam(p, lr=0.01)
    new = opt.step(p, g)[0]
    # first bias-corrected step: m_hat = g, v_hat = g^2 -> p - lr * g / (|g| + eps)
    assert np.allclose(new, p[0] - 0.01 * g[0] / (np.abs(g[0]) + 1e-8))


# ------------------------------------------------------------------ GPU kernels

def test_fp16_oracle_is_close_to_the_exact_one():
    """The FP16-rounded restatement against the float64 model: outputs within half
    precision, and every layer's gradient within 1e-2 when the float64 backward uses the
    same ReLU masks.  (Rounding the activations also flips ~1e-4 of the masks; records whose
    mean-direction output is near zero carry gradients ~1/|m| large, so such flips move the
    summed gradients by up to ~20 % -- a property of the half-precision model, which the
    kernels share, not of the restatement.)"""
    Ws, bs = G.init_params(4)
    x, d, c = _batch(512, 9)
    X = G.encode(x)
    o64, acts, pre = G.forward(Ws, bs, X)
    o16, acts16 = G.forward_f16(Ws, bs, x)
    assert np.abs(o16 - o64).max() < 2e-2 * np.abs(o64).max()
    loss16, dW16, db16 = G.grads_f16(Ws, bs, x, d, c)
    _, dout = G.loss_and_grad_out(o64, d, c)
    # float64 backward with the FP16 forward's masks (acts16[i] > 0 iff its pre-activation was)
    dWs, dbs = G.backward(Ws, bs, acts, acts16[1:], dout)
    for l in range(6):
        a, b = dW16[l], dWs[l][:, :G.N_FEAT] if l == 0 else dWs[l]
        assert np.linalg.norm(a - b) < 1e-2 * np.linalg.norm(b), l
        assert np.linalg.norm(db16[l] - dbs[l]) < 1e-2 * np.linalg.norm(dbs[l]), l


@pytest.mark.gpu
def test_gpu_guide_forward_matches_oracle(tracer):
    """inference against the FP16-rounded restatement (tight) and the float64 model (the
    rounding's own size)"""
    Ws, bs = G.init_params(7)
    tracer.guide_set_params(G.to_padded(Ws, bs))
    x, _, _ = _batch(1000, 1)
    got = tracer.guide_forward(x)[:, :40]
    ref16, _ = G.forward_f16(Ws, bs, x)
    ref, _, _ = G.forward(Ws, bs, G.encode(x))
    scale = np.abs(ref).max()
    err16 = np.abs(got - ref16).max() / scale
    err = np.abs(got - ref).max() / scale
    print("forward max error / output scale: vs FP16 oracle", err16, "vs float64", err)
    assert err16 < 2e-3
    assert err < 2e-2
    assert np.corrcoef(got.ravel(), ref.ravel())[0, 1] > 0.9995


@pytest.mark.gpu
@pytest.mark.parametrize("n", [4096, 1000])
def test_gpu_guide_gradients_match_oracle(tracer, n):
    """the fused training step's loss and mean gradients of every layer against the
    FP16-rounded restatement (relative error <= 1e-2 per layer; 1000 rows: a ragged last
    chunk) and, loosely, the float64 model"""
    Ws, bs = G.init_params(11)
    tracer.guide_set_params(G.to_padded(Ws, bs))
    x, d, c = _batch(n, 2)
    loss, grad = tracer.guide_train(x, d, c, apply=False, want_grad=True)
    loss16, dW16, db16 = G.grads_f16(Ws, bs, x, d, c)
    X = G.encode(x)
    o, acts, pre = G.forward(Ws, bs, X)
    ref_loss, dout = G.loss_and_grad_out(o, d, c)
    dWs, dbs = G.backward(Ws, bs, acts, pre, dout)
    gW, gb = G.from_padded(grad)
    print("loss", loss, "FP16 oracle", loss16, "float64", ref_loss)
    assert abs(loss - loss16) < 1e-3 * abs(loss16) + 1e-4
    assert abs(loss - ref_loss) < 1e-2 * abs(ref_loss) + 1e-3
    for l in range(6):
        pairs = ((gW[l][:, :G.N_FEAT] if l == 0 else gW[l], dW16[l], dWs[l][:, :G.N_FEAT] if l == 0 else dWs[l], "W"),
                 (gb[l], db16[l], dbs[l], "b"))
        for a, b16, b64, name in pairs:
            rel16 = np.linalg.norm(a - b16) / (np.linalg.norm(b16) + 1e-30)
            rel = np.linalg.norm(a - b64) / (np.linalg.norm(b64) + 1e-30)
            print(f"layer {l} {name}: relative error vs FP16 oracle {rel16:.2e}, vs float64 {rel:.2e}")
            assert rel16 < 1e-2, (l, name, rel16)
            assert rel < 0.3, (l, name, rel)  # ReLU masks flipped by the half rounding


@pytest.mark.gpu
def test_gpu_adam_step(tracer):
    Ws, bs = G.init_params(3)
    p0 = G.to_padded(Ws, bs)
    tracer.guide_set_params(p0)
    x, d, c = _batch(2048, 4)
    _, grad = tracer.guide_train(x, d, c, apply=False, want_grad=True)
    tracer.guide_set_params(p0)          # clears the Adam state
    tracer.guide_train(x, d, c, lr=1e-3, apply=True)
    p1 = tracer.guide_params()
    live = np.abs(grad) > 1e-5
    # first step: p1 = p0 - lr * g / (|g| + eps)
    assert np.allclose((p1 - p0)[live], -1e-3 * np.sign(grad[live]), atol=2e-5)


@pytest.mark.gpu
def test_gpu_guide_kernel_time_and_loss_free_step(tracer):
    """rbw_guide_last_ms reports the network kernels of the last call; a step without the
    loss read-back updates the parameters as one with it"""
    Ws, bs = G.init_params(5)
    p0 = G.to_padded(Ws, bs)
    x, d, c = _batch(4096, 6)
    tracer.guide_set_params(p0)
    tracer.guide_train(x, d, c, lr=1e-3, apply=True)
    ms_train = tracer.guide_last_ms()
    p_with = tracer.guide_params()
    tracer.guide_set_params(p0)
    loss, _ = tracer.guide_train(x, d, c, lr=1e-3, apply=True, want_loss=False)
    p_without = tracer.guide_params()
    assert loss is None
    # (float atomics in the batch sums: equal up to their summation order)
    assert np.allclose(p_with, p_without, rtol=0, atol=1e-6)
    tracer.guide_forward(x)
    ms_fwd = tracer.guide_last_ms()
    print("train step", ms_train, "ms, inference", ms_fwd, "ms for", len(x), "records")
    assert 0.0 < ms_train < 1000.0 and 0.0 < ms_fwd < 1000.0


@pytest.mark.gpu
def test_gpu_training_learns_a_direction(tracer):
    """samples whose radiance comes from one direction (depending on the input) are learned"""
    tracer.guide_reset(5)
    rng = np.random.default_rng(9)
    target = lambda x: np.stack([np.cos(4 * x[:, 0]), np.sin(4 * x[:, 0]), 0.3 + 0 * x[:, 0]], 1)
    first = last = None
    for it in range(150):
        x = rng.random((8192, 9))
        t = target(x)
        t /= np.linalg.norm(t, axis=1, keepdims=True)
        d = rng.normal(size=(8192, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        # radiance around the target direction, sampled uniformly on the sphere (pdf 1/4pi)
        L = np.exp(20.0 * ((d * t).sum(1) - 1.0))
        loss, _ = tracer.guide_train(x, d, L * 4 * math.pi, lr=5e-3)
        first = loss if first is None else first
        last = loss
    x = rng.random((2000, 9))
    t = target(x)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    q_t = G.mixture_pdf(tracer.guide_forward(x)[:, :40].astype(np.float64), t)
    print("loss", first, "->", last, "median q(target)", np.median(q_t))
    assert last < first - 1.0
    assert np.median(q_t) > 1.0          # uniform density is 1/(4 pi) = 0.08


@pytest.mark.gpu
@pytest.mark.parametrize("guide_depth", [0, 1, 2])
def test_gpu_guided_render_unbiased(tracer, guide_depth):
    """the guided estimator's expectation equals the plain one (Cornell box, mean radiance),
    also when only the first vertices are guided and later ones shade as the plain estimator"""
    from rainbow_amd.tracer import guiding_opts, render_params
    sc, cam = scenes.cornell_box()
    tracer.upload(sc)
    tracer.guide_reset(1)
    plain = tracer.render(cam, render_params(spp=64, max_depth=5, seed=21))["image"]
    go = guiding_opts(train=True, steps_per_pass=2, batch=1 << 16, guide_depth=guide_depth)
    guided = tracer.render(cam, render_params(spp=64, max_depth=5, seed=22, guiding=C_ref(go)))
    gi = guided["image"]
    assert np.isfinite(gi).all()
    mp, mg = plain.mean(axis=(0, 1)), gi.mean(axis=(0, 1))
    print("plain", mp, "guided", mg, guided["stats"]["kernel_ms"])
    assert np.all(np.abs(mg - mp) < 0.01 * mp + 1e-4)


@pytest.mark.gpu
def test_gpu_guided_furnace(tracer):
    """closed emissive box: E[L] = Le * sum_k rho^k also when the continuation is guided"""
    from rainbow_amd.tracer import guiding_opts, render_params
    from tests.test_oracle import closed_box, furnace_expected, FURNACE_DEPTH
    sc, cam = closed_box(0.5, 1.0)
    tracer.upload(sc)
    tracer.guide_reset(2)
    for gd in (0, 2):
        go = guiding_opts(train=True, steps_per_pass=2, batch=1 << 14, guide_depth=gd)
        r = tracer.render(cam, render_params(spp=256, max_depth=FURNACE_DEPTH, seed=4, guiding=C_ref(go)))
        mean = float(r["image"].mean())
        expect = furnace_expected(0.5, 1.0, FURNACE_DEPTH)
        assert abs(mean - expect) < 0.01 * expect, (gd, mean, expect)
    go = guiding_opts(train=True, steps_per_pass=2, batch=1 << 14)
    r = tracer.render(cam, render_params(spp=256, max_depth=FURNACE_DEPTH, seed=4, guiding=C_ref(go)))
    mean = float(r["image"].mean())
    expect = furnace_expected(0.5, 1.0, FURNACE_DEPTH)
    assert abs(mean - expect) < 0.01 * expect, (mean, expect)


@pytest.mark.gpu
def test_gpu_trained_guide_reduces_error(tracer):
    """What the trained network gains: on the narrow-aperture room (beam through the gap onto
    the ceiling, the view lit indirectly) the guided estimator, after online training,
    has lower relMSE than the plain one at equal spp, with the same expectation."""
    from rainbow_amd.tracer import guiding_opts, render_params
    sc, cam = scenes.narrow_aperture_room(level=5, width=96, height=96, light_tilt=45.0)
    tracer.upload(sc)
    ref = tracer.render(cam, render_params(spp=16384, max_depth=8, seed=100))["image"]
    plain = tracer.render(cam, render_params(spp=256, max_depth=8, seed=200))["image"]
    tracer.guide_reset(3)
    for i in range(16):
        go = guiding_opts(train=True, steps_per_pass=4, batch=1 << 16)
        tracer.render(cam, render_params(spp=16, max_depth=8, seed=300 + i, guiding=C_ref(go)))
    go = guiding_opts(train=False)
    guided = tracer.render(cam, render_params(spp=256, max_depth=8, seed=400, guiding=C_ref(go)))["image"]
    ep, eg = scenes.rel_mse(plain, ref), scenes.rel_mse(guided, ref)
    print("relMSE plain", ep, "guided", eg, "ratio", eg / ep)
    assert abs(guided.mean() - ref.mean()) < 0.01 * ref.mean()
    assert eg < 0.97 * ep


def C_ref(opts):
    import ctypes as C
    return C.pointer(opts)


@pytest.fixture(scope="module")
def room512(tracer):
    """the narrow-aperture room at its configured size (level-8 mesh, 512x512, depth 8, the
    beam through the 0.3 x 0.3 gap onto the ceiling) and a plain judge of 8192 spp"""
    from rainbow_amd.tracer import render_params
    sc, cam = scenes.narrow_aperture_room(level=8, width=512, height=512, light_tilt=45.0)
    tracer.upload(sc)
    judge = tracer.render(cam, render_params(spp=8192, max_depth=8, seed=1 << 20), stats=False)["image"]
    return sc, cam, judge


def _chunked(tracer, cam, chunks, spp, seed0, **kw):
    """the average of `chunks` independent renders and the standard error of its image mean"""
    from rainbow_amd.tracer import render_params
    imgs = [tracer.render(cam, render_params(spp=spp, max_depth=8, seed=seed0 + k, **kw), stats=False)["image"]
            for k in range(chunks)]
    means = np.array([float(im.mean()) for im in imgs])
    return np.mean(imgs, axis=0), float(means.std(ddof=1) / np.sqrt(chunks))


@pytest.mark.gpu
@pytest.mark.parametrize("guide_depth,bar", [(1, 0.97), (0, 0.95)])
def test_gpu_guided_room_at_the_configured_size(tracer, room512, guide_depth, bar):
    """The learned-guiding workload at its configured size: the network trained online from
    scratch with the benchmark's schedule (32 renders of 8 spp, 4 Adam steps of 2^18 records
    after each), then 8 independent 32-spp renders of the guided (vertex 0, or every vertex)
    and of the plain estimator.  The guided image mean equals the judge's within 4 standard
    errors (unbiased), and its relMSE at equal spp is below `bar` times the plain one's."""
    from rainbow_amd.tracer import guiding_opts
    sc, cam, judge = room512
    tracer.upload(sc)
    tracer.guide_reset(7)
    for i in range(32):
        go = guiding_opts(train=True, steps_per_pass=4, batch=1 << 18, guide_depth=guide_depth)
        tracer.render(cam, render_params_guided(i, go), stats=False)
    go = guiding_opts(train=False, guide_depth=guide_depth)
    guided, se_g = _chunked(tracer, cam, 8, 32, 5000, guiding=C_ref(go))
    plain, se_p = _chunked(tracer, cam, 8, 32, 9000)
    mj, mg, mp = float(judge.mean()), float(guided.mean()), float(plain.mean())
    eg, ep = scenes.rel_mse(guided, judge), scenes.rel_mse(plain, judge)
    print(f"guide_depth {guide_depth}: means judge {mj:.5f} guided {mg:.5f} (se {se_g:.1e}) plain {mp:.5f} "
          f"(se {se_p:.1e}); relMSE guided {eg:.3e} plain {ep:.3e} ratio {eg / ep:.3f}")
    assert abs(mg - mj) < 4 * se_g * np.sqrt(1 + 256 / 8192)
    assert abs(mp - mj) < 4 * se_p * np.sqrt(1 + 256 / 8192)
    assert eg < bar * ep


def render_params_guided(i, go):
    from rainbow_amd.tracer import render_params
    return render_params(spp=8, max_depth=8, seed=300 + i, guiding=C_ref(go))
