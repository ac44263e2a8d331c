"""Closed-form pins of three oracle functions that GPU-vs-oracle parity alone
cannot fix (both sides would share a mistake):

* camera ray generation, perspective_camera::sample (rainbow-cpu/cameras/
  perspective_camera.cpp:78-99) with the projective camera's matrices
  (cameras/camera.cpp:21-52): pinhole rays against pinhole geometry (view
  direction at the film centre, corners at +-tan(fov/2) with raster y pointing
  down), and thin-lens rays through one focus point per raster position;
* sphere intersection, sphere::intersect (shapes/sphere.cpp:17-38) with the
  double-precision quadratic (rainbow-core/utilities.cpp:7-31): distances along
  rays aimed at the centre, the far root from inside, the tangent boundary, and
  hit points on the radius;
* Russian roulette, path_integrator::trace (integrators/path_integrator.cpp:35-43):
  furnaces deep enough for roulette to fire, against the expectation
  Le * sum_k rho^k (unbiasedness: a wrong survival weight fails it) and against
  the number of extension rays the survival rule implies (a wrong q, threshold
  comparison, start bounce or 0.05 floor fails it).

The helpers are shared with the GPU versions in tests/test_gpu_pins.py.
"""
import math

import numpy as np
import pytest

from rainbow_amd import abi, scenes


def _orc():
    from oracle import orc_py
    return orc_py


# ------------------------------------------------------------------ camera

def camera_basis(eye, target, up):
    """world axes of a right-handed look-at camera: right, up, forward"""
    e, t, u = (np.asarray(v, np.float64) for v in (eye, target, up))
    f = (t - e) / np.linalg.norm(t - e)
    r = np.cross(f, u)
    r /= np.linalg.norm(r)
    return r, np.cross(r, f), f


def pinhole_directions(xy, width, height, fov_deg, eye, target, up=(0, 1, 0)):
    """world directions of a pinhole camera with square pixels: raster (x, y), y down,
    maps to the image-plane point ((2x/W - 1) t W/H, (1 - 2y/H) t) at distance 1 along
    the view direction, t = tan(fov/2)"""
    r, u, f = camera_basis(eye, target, up)
    t = math.tan(math.radians(fov_deg) / 2)
    px = (2 * xy[:, 0] / width - 1) * t * width / height
    py = (1 - 2 * xy[:, 1] / height) * t
    d = px[:, None] * r + py[:, None] * u + f
    return d / np.linalg.norm(d, axis=1, keepdims=True)


@pytest.mark.parametrize("size", [(64, 64), (96, 48), (40, 80)])
def test_oracle_pinhole_rays_match_camera_geometry(size):
    W, H = size
    eye, target, up = (0.3, -0.2, 3.0), (0.1, 0.2, -0.5), (0.1, 1.0, 0.0)
    cam = scenes.perspective_camera(W, H, 40.0, eye=eye, target=target, up=up)
    rng = np.random.default_rng(1)
    xy = np.concatenate([rng.uniform(0, 1, (500, 2)) * (W, H),
                         [[W / 2, H / 2], [0, 0], [W, H], [W, 0], [0, H]]]).astype(np.float32)
    o, d = _orc().camera_rays(cam, xy)
    want = pinhole_directions(xy.astype(np.float64), W, H, 40.0, eye, target, up)
    # chord between unit vectors (= the angle for small angles; arccos loses it near 1)
    ang = np.linalg.norm(d.astype(np.float64) - want, axis=1)
    print("max angle to pinhole geometry", ang.max())
    assert ang.max() < 1e-6
    assert np.abs(o - np.asarray(eye, np.float32)).max() < 1e-6
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-6
    # the film centre looks along the view direction; the top-left corner is up and to
    # the left of it at tan(fov/2) vertically (raster y points down the film)
    r, u, f = camera_basis(eye, target, up)
    c, tl, br = d[-5].astype(np.float64), d[-4].astype(np.float64), d[-3].astype(np.float64)
    assert np.linalg.norm(c - f) < 1e-6
    t = math.tan(math.radians(20.0))
    for corner, sx, sy in ((tl, -1, 1), (br, 1, -1)):
        plane = corner / np.dot(corner, f)
        assert abs(np.dot(plane, u) - sy * t) < 1e-5
        assert abs(np.dot(plane, r) - sx * t * W / H) < 1e-5


def test_oracle_thin_lens_rays_meet_at_the_focus_point():
    """Every lens sample of one raster position gives a ray through the same point: the
    pinhole ray's point at camera-space z = focus_distance (perspective_camera.cpp:88-97;
    a right-handed camera looks down -z, so a focus plane in front of the lens has a
    negative focus_distance).  Origins lie on the lens disk of the given radius."""
    W, H, fov, lens, focus = 48, 32, 35.0, 0.12, -2.5
    eye, target = (0.2, 0.1, 3.0), (0.0, -0.1, 0.0)
    cam = scenes.perspective_camera(W, H, fov, eye=eye, target=target, lens_radius=lens, focus=focus)
    r, u, f = camera_basis(eye, target, (0, 1, 0))
    rng = np.random.default_rng(5)
    for xy0 in ([W / 2, H / 2], [3.25, 7.5], [W - 0.5, H - 0.25]):
        xy = np.tile(np.asarray(xy0, np.float32), (64, 1))
        uv = rng.uniform(0, 1, (64, 2)).astype(np.float32)
        o, d = _orc().camera_rays(cam, xy, uv)
        pin = pinhole_directions(np.asarray([xy0], np.float64), W, H, fov, eye, target)[0]
        # camera-space z of the pinhole direction is -dot(pin, f): scale it to z = focus
        pf = np.asarray(eye, np.float64) + pin * (focus / -np.dot(pin, f))
        o64, d64 = o.astype(np.float64), d.astype(np.float64)
        # distance from the focus point to each ray's line, and the point lies ahead of the origin
        v = pf - o64
        along = np.sum(v * d64, axis=1)
        miss = np.linalg.norm(v - along[:, None] * d64, axis=1)
        print("focus miss", miss.max())
        assert miss.max() < 5e-6 * np.linalg.norm(v, axis=1).max()
        assert (along > 0).all()
        off = o64 - np.asarray(eye, np.float64)
        assert np.abs(off @ f).max() < 1e-6  # the lens is the camera's z = 0 plane
        assert np.linalg.norm(off, axis=1).max() <= lens * (1 + 1e-6)
        assert np.linalg.norm(off, axis=1).max() > 0.3 * lens  # the samples spread over the lens
    # the lens centre (concentric map of (0.5, 0.5)) gives the pinhole ray
    xy = np.asarray([[10.5, 20.25]], np.float32)
    o, d = _orc().camera_rays(cam, xy, np.asarray([[0.5, 0.5]], np.float32))
    pin = pinhole_directions(xy.astype(np.float64), W, H, fov, eye, target)[0]
    assert np.linalg.norm(d[0].astype(np.float64) - pin) < 1e-6


# ------------------------------------------------------------------ sphere

SPHERE_C = np.array([0.3, -0.2, 0.1])
SPHERE_R = 0.45
SPHERE_PRIM = 2  # primitive id of the sphere: spheres follow the scene's two triangles


def sphere_scene():
    sc = scenes.Scene()
    sc.add_sphere(tuple(SPHERE_C), SPHERE_R, sc.matte((0.5, 0.5, 0.5)))
    # a far emitter keeps the scene renderable; it lies outside every test ray's reach
    P, I, N = scenes._quad((-1, 50, -1), (1, 50, -1), (1, 50, 1), (-1, 50, 1))
    sc.add_mesh(P, I, sc.matte((0, 0, 0)), normals=-N, radiance=(1.0, 1.0, 1.0))
    return sc


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def sphere_ray_sets(seed=3, n=4000):
    """(name, origins, directions, expected t or None for a miss) in float32; the
    expected distances are exact in float64 for the float32 inputs"""
    rng = np.random.default_rng(seed)
    c, r = SPHERE_C, SPHERE_R
    out = []
    # aimed at the centre from outside: t = |c - o| - r
    u = _unit(rng.normal(size=(n, 3)))
    o = (c + u * rng.uniform(0.6, 5.0, (n, 1))[:, :1]).astype(np.float32)
    d = _unit(c - o.astype(np.float64)).astype(np.float32)
    o64, d64 = o.astype(np.float64), d.astype(np.float64)
    b = np.sum((o64 - c) * d64, axis=1)
    cc = np.sum((o64 - c) ** 2, axis=1) - r * r
    out.append(("aimed at the centre", o, d, -b - np.sqrt(b * b - cc)))
    # from inside: the far root t1 = -b + sqrt(b^2 - c)
    o = (c + _unit(rng.normal(size=(n, 3))) * rng.uniform(0.0, 0.9, (n, 1)) * r).astype(np.float32)
    d = _unit(rng.normal(size=(n, 3))).astype(np.float32)
    o64, d64 = o.astype(np.float64), d.astype(np.float64)
    b = np.sum((o64 - c) * d64, axis=1)
    cc = np.sum((o64 - c) ** 2, axis=1) - r * r
    out.append(("from inside", o, d, -b + np.sqrt(b * b - cc)))
    # near-tangent: impact parameter r (1 -+ 1e-3) hits / misses
    for name, k in (("just inside the tangent", 1 - 1e-3), ("just outside the tangent", 1 + 1e-3)):
        d = _unit(rng.normal(size=(n, 3)))
        side = _unit(np.cross(d, rng.normal(size=(n, 3))))
        o = (c + side * r * k - d * rng.uniform(1.0, 3.0, (n, 1))).astype(np.float32)
        d = d.astype(np.float32)
        o64, d64 = o.astype(np.float64), d.astype(np.float64)
        b = np.sum((o64 - c) * d64, axis=1)
        cc = np.sum((o64 - c) ** 2, axis=1) - r * r
        disc = b * b - cc
        out.append((name, o, d, (-b - np.sqrt(np.maximum(disc, 0))) if k < 1 else None))
    return out


def check_sphere_hits(name, o, d, want, hits):
    """hits: rbw_hit records of the rays (entity-local t of a sphere translated to its
    centre equals the world distance along the unit direction up to rounding)"""
    if want is None:
        assert (hits["prim"] != SPHERE_PRIM).all(), name
        return
    assert (hits["prim"] == SPHERE_PRIM).all(), (name, np.count_nonzero(hits["prim"] != SPHERE_PRIM))
    t = hits["t"].astype(np.float64)
    # the quadratic's coefficients are formed in float32 (sphere.cpp:20-24): c = o.o - r^2
    # carries a few ulps of |o - centre|^2, which move a root by about that over
    # 2 sqrt(discriminant) (r for a ray through the centre, ~0.045 r at 1e-3 from the
    # tangent).  A wrong root or sign is off by the order of r.
    o64, d64 = o.astype(np.float64), d.astype(np.float64)
    dist2 = np.sum((o64 - SPHERE_C) ** 2, axis=1)
    b = np.sum((o64 - SPHERE_C) * d64, axis=1)
    sq = np.sqrt(np.maximum(b * b - (dist2 - SPHERE_R ** 2), 0.0))
    tol = 8 * 2.0 ** -24 * (dist2 + SPHERE_R ** 2) / (2 * sq) + 1e-7
    err = np.abs(t - want) / tol
    print(name, "max t error / float32 bound", err.max())
    assert err.max() < 1, name
    p = o.astype(np.float64) + t[:, None] * d.astype(np.float64)
    rad = np.linalg.norm(p - SPHERE_C, axis=1)
    assert (np.abs(rad - SPHERE_R) < tol + 1e-7 * np.sqrt(dist2)).all(), name


def test_oracle_sphere_closed_forms():
    os_ = _orc().OracleScene(sphere_scene())
    for name, o, d, want in sphere_ray_sets():
        hits = os_.trace(o, d, np.full(len(o), np.inf, np.float32))
        check_sphere_hits(name, o, d, want, hits)


def test_oracle_sphere_respects_the_ray_length():
    """a hit beyond the ray's length is no hit (sphere.cpp:27: t0 > ray.length)"""
    os_ = _orc().OracleScene(sphere_scene())
    name, o, d, want = sphere_ray_sets(seed=9, n=500)[0]
    short = os_.trace(o, d, (want * 0.999).astype(np.float32))
    assert (short["prim"] != SPHERE_PRIM).all()
    long_ = os_.trace(o, d, (want * 1.001 + 1e-6).astype(np.float32))
    assert (long_["prim"] == SPHERE_PRIM).all()


# ------------------------------------------------------------------ Russian roulette

# (albedo, RR threshold, max depth): roulette from bounce 4 at every vertex (0.5, 1);
# skipped at vertices whose throughput is back above the threshold (0.8, 0.5); the 0.05
# floor of q for throughputs at or above 0.95 (0.97, 2)
RR_CASES = [(0.5, 1.0, 10), (0.8, 0.5, 12), (0.97, 2.0, 12)]


def furnace_expected(albedo, Le, depth):
    """emission at the first vertex plus next-event estimation at vertices 0..depth-1"""
    return Le * sum(albedo ** k for k in range(depth + 1))


def rr_survival(albedo, threshold, depth):
    """P(a path traces bounce b), b = 0..depth-1, in the closed emissive Lambert box:
    cosine-sampled continuations multiply the throughput by exactly the albedo, so the
    throughput, and with it each roulette's q = max(0.05, 1 - max(beta)) for bounces > 3
    below the threshold, is the same for every path (path_integrator.cpp:35-43)"""
    alive, beta, out = 1.0, 1.0, []
    for b in range(depth):
        out.append(alive)
        beta *= albedo
        if beta < threshold and b > 3:
            q = max(0.05, 1.0 - beta)
            alive *= 1.0 - q
            beta /= 1.0 - q
    return np.array(out)


def rr_ray_moments(albedo, threshold, depth):
    """mean and variance of the extension rays of one camera sample"""
    s = rr_survival(albedo, threshold, depth)
    b = np.arange(depth)
    mean = s.sum()
    return mean, (s * (2 * b + 1)).sum() - mean * mean


def check_furnace(image, rays, samples, albedo, threshold, depth, check_rays=True):
    expect = furnace_expected(albedo, 1.0, depth)
    px = image[..., 0].astype(np.float64).ravel()
    mean, se = px.mean(), px.std() / math.sqrt(px.size)
    print(f"furnace rho {albedo} thr {threshold} depth {depth}: mean {mean:.5f} expect {expect:.5f} "
          f"se {se:.2e} rays/sample {rays / samples:.4f}")
    assert abs(mean - expect) < 5 * se + 1e-4 * expect, (mean, expect, se)
    # the bar resolves the part of the expectation that roulette weights
    assert 5 * se < 0.5 * sum(albedo ** k for k in range(6, depth + 1))
    if check_rays:
        m, v = rr_ray_moments(albedo, threshold, depth)
        got = rays / samples
        assert abs(got - m) < 5 * math.sqrt(v / samples) + 1e-6, (got, m)


def closed_box(albedo, Le):
    from test_oracle import closed_box as cb
    return cb(albedo, Le)


def test_rr_survival_reduces_to_no_roulette_at_shallow_depth():
    # roulette starts after bounce 3: depth 4 traces every bounce of every path
    assert np.array_equal(rr_survival(0.5, 1.0, 4), np.ones(4))
    assert rr_ray_moments(0.5, 1.0, 4) == (4.0, 0.0)


@pytest.mark.parametrize("albedo,threshold,depth", RR_CASES)
def test_oracle_furnace_with_russian_roulette(albedo, threshold, depth):
    sc, cam = closed_box(albedo, 1.0)
    orc = _orc()
    r = orc.OracleScene(sc).render(cam, orc.render_params(spp=256, max_depth=depth, rr_threshold=threshold, seed=7),
                                   sampler=orc.ORC_SAMPLER_PER_SAMPLE)
    st = r["stats"]
    check_furnace(r["image"], st["extension_rays"], st["samples"], albedo, threshold, depth)


# ------------------------------------------------------------------ fexp near its overflow

def test_fexp_is_finite_up_to_the_float_limit():
    import ctypes as C
    L = _orc().lib()
    L.orc_fexp.argtypes = [C.c_float]
    L.orc_fexp.restype = C.c_float
    xs = np.linspace(88.0, 88.7228, 200, dtype=np.float32)
    got = np.array([L.orc_fexp(float(x)) for x in xs], np.float32)
    assert np.isfinite(got).all()
    want = np.exp(xs.astype(np.float64))
    ulp = np.spacing(want.astype(np.float32)).astype(np.float64)
    assert (np.abs(got.astype(np.float64) - want) <= 2 * ulp).all()
    assert L.orc_fexp(88.5) > 2.7e38
