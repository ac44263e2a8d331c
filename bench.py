"""Benchmark of the path-integrator render call on one MI355X.

A step is one full render of the workload (every camera sample of the film
traced to completion and accumulated), inputs resident in HBM.  Default
workload: an fBm-displaced icosphere subdivided 8 times (1,310,720 triangles,
per-triangle Lambert or GGX-conductor materials) inside the Cornell box,
1024x1024, 64 spp, max depth 8, Russian roulette from depth 3, seed 0.

Prints one JSON line: Msamples/s (whole job), Mrays/s, the roofline of the
dominant kernel (algorithmic bytes / HIP-event time on the library's stream)
and the CPU baseline (the oracle restatement of rainbow-cpu's path integrator,
reference per-tile pcg32 stream, a bounded subset of tiles on the host cores).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s + time to relMSE 1e-3: guided vs plain GPU PT, 1 MI355X"


_NATIVE_BUILD = None  # the child process building liborc_native.so (main starts it)


def _join_native_build():
    global _NATIVE_BUILD
    if _NATIVE_BUILD is None:
        return
    err = _NATIVE_BUILD.communicate()[1]
    if _NATIVE_BUILD.returncode != 0:
        print(f"# native oracle build failed ({_NATIVE_BUILD.returncode}): {err.decode()[-400:]}; timing the "
              "parity build", file=sys.stderr, flush=True)
    _NATIVE_BUILD = None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(sc, cam, spp, depth, target_s=15.0, one_core_s=None, **pkw):
    """The oracle on a bounded subset of tiles (every `stride`-th 16x16 tile), all spp, on all
    the host threads the box gives (at most 16) and on one: the timing build
    (liborc_native.so, -O3 -march=native for this host) when bench.py built it, else the
    parity build.  The timed region is the oracle's tile loop and film merge, the
    reference's (sampler_integrator.cpp:90-143)."""
    from oracle import orc_py  # the CPU restatement, timed as the baseline
    _join_native_build()
    threads = min(16, os.cpu_count() or 1)
    native = orc_py.native_lib() is not None
    os_ = orc_py.OracleScene(sc, native=native)
    p = orc_py.render_params(spp=spp, max_depth=depth, **pkw)
    ntiles = ((cam.width + 15) // 16) * ((cam.height + 15) // 16)

    def timed(nthreads, seconds):
        stride = max(1, ntiles // (8 * nthreads))
        r = os_.render(cam, p, sampler=orc_py.ORC_SAMPLER_TILE, threads=nthreads, tile_stride=stride)
        rate = r["stats"]["samples"] / max(r["stats"]["seconds"], 1e-9)
        want_tiles = max(1.0, rate * seconds / (256 * spp))
        stride = max(1, int(ntiles / want_tiles))
        st = os_.render(cam, p, sampler=orc_py.ORC_SAMPLER_TILE, threads=nthreads, tile_stride=stride)["stats"]
        return st, stride

    st, stride = timed(threads, target_s)
    st1, stride1 = timed(1, one_core_s if one_core_s else target_s / 3)
    return {
        "value": st["samples"] / st["seconds"] / 1e6,
        "unit": "Msamples/s",
        "cores": st["threads"],
        "kind": "port",
        "value_1core": st1["samples"] / st1["seconds"] / 1e6,
        "cpu_model": cpu_model(),
        "build": "g++ -O3 -march=native (built on this host)" if native else "g++ -O2 -ffp-contract=off (parity build)",
        "sample": f"every {stride}th 16x16 tile of the same workload at full spp/depth: "
                  f"{st['samples']} samples in {st['seconds']:.2f} s on {st['threads']} threads; 1 core: every "
                  f"{stride1}th tile, {st1['samples']} samples in {st1['seconds']:.2f} s",
    }


def _time_to(curve, threshold):
    """first crossing of relMSE <= threshold on a (ms, relMSE) curve, log-log interpolated"""
    import math
    for (t0, e0), (t1, e1) in zip(curve, curve[1:]):
        if e1 <= threshold < e0:
            f = (math.log(e0) - math.log(threshold)) / (math.log(e0) - math.log(e1))
            return t0 + f * (t1 - t0)
    if curve and curve[0][1] <= threshold:
        return curve[0][0]
    return None


def _err_at(curve, t_ms):
    """relMSE at time t_ms on a (ms, relMSE) curve, log-log interpolated (None past its end)"""
    import math
    if not curve or t_ms < curve[0][0]:
        return None
    for (t0, e0), (t1, e1) in zip(curve, curve[1:]):
        if t0 <= t_ms <= t1:
            f = (math.log(t_ms) - math.log(t0)) / (math.log(t1) - math.log(t0))
            return math.exp(math.log(e0) + f * (math.log(e1) - math.log(e0)))
    return curve[-1][1] if abs(t_ms - curve[-1][0]) < 1e-9 else None


def guided_leg(full, threshold=1e-3, guide_depth=1, product_targets=True, analysis=False):
    """Time to relMSE <= threshold, plain vs guided, on the narrow-aperture room.

    The room is lit only through a 0.3x0.3 gap in its +x wall by a 0.5x0.5 quad light
    (radiance 50) outside, aimed 45 degrees upward through the gap, so the beam lands on
    the ceiling and the camera's view is lit indirectly.  The judge is a plain render of
    judge_spp samples per pixel; time is the sum of the render calls' wall time (all
    kernels: tracing, MLP inference, training), relMSE computed between calls.  The guided
    estimator guides the first guide_depth vertices of each path and shades later ones as
    the plain tracer (0: every vertex); the leg also reports the every-vertex schedule
    against the same judge.  product_targets: the network is trained on f |cos| L_i (else
    L_i).  analysis (full size only): also the offline-fitted guide's ceiling and the
    variance by path depth (several minutes more).
    """
    import numpy as np
    from rainbow_amd import scenes
    from rainbow_amd.tracer import Tracer, guiding_opts, render_params
    res, judge_spp = (512, 65536) if full else (256, 16384)
    sc, cam = scenes.narrow_aperture_room(level=8, width=res, height=res, light_tilt=45.0)
    tr = Tracer(0)
    tr.upload(sc)
    depth, chunk, max_spp, train_calls = 8, 8, 4096, 32
    t0 = time.perf_counter()
    ref = tr.render(cam, render_params(spp=judge_spp, max_depth=depth, seed=1 << 20))["image"]
    judge_s = time.perf_counter() - t0
    print(f"# guided leg: judge {judge_spp} spp in {judge_s:.1f} s", file=sys.stderr, flush=True)

    def run(guided, gd=0):
        acc = np.zeros_like(ref)
        elapsed, curve, n = 0.0, [], 0
        if guided:
            tr.guide_reset(7)
        while n * chunk < max_spp:
            kw = {}
            if guided:
                go = guiding_opts(train=n < train_calls, steps_per_pass=4, batch=1 << 18, guide_depth=gd,
                                  product_targets=product_targets)
                kw["guiding"] = C.pointer(go)
            t = time.perf_counter()
            img = tr.render(cam, render_params(spp=chunk, max_depth=depth, seed=(2 if guided else 1) * 100000 + n,
                                               **kw), stats=False)["image"]
            elapsed += time.perf_counter() - t
            acc += img
            n += 1
            e = scenes.rel_mse(acc / n, ref)
            curve.append((elapsed * 1e3, e))
            if e <= threshold * 0.8:
                break
        return curve, n * chunk

    pc, pspp = run(False)
    gc, gspp = run(True, guide_depth)
    tp, tg = _time_to(pc, threshold), _time_to(gc, threshold)
    ac, aspp = run(True, 0) if guide_depth != 0 else (gc, gspp)
    ta = _time_to(ac, threshold)

    # The ceiling of this guide: a network fitted offline (800 Adam steps of 2^18 records from
    # 1600 spp of training renders, learning rate 1e-2 then 3e-3 then 1e-3), frozen, timed
    # without its training against the same judge.
    def offline(gd, alpha):
        tr.guide_reset(11)
        for n in range(200):
            lr = 1e-2 if n < 100 else (3e-3 if n < 150 else 1e-3)
            go = guiding_opts(train=True, steps_per_pass=4, batch=1 << 18, guide_depth=gd, learning_rate=lr,
                              product_targets=product_targets)
            tr.render(cam, render_params(spp=8, max_depth=depth, seed=7 * 100000 + n, guiding=C.pointer(go)),
                      stats=False)
        acc = np.zeros_like(ref)
        elapsed, curve, n = 0.0, [], 0
        while n * chunk < max_spp:
            go = guiding_opts(mis_fraction=alpha, train=False, guide_depth=gd)
            t = time.perf_counter()
            img = tr.render(cam, render_params(spp=chunk, max_depth=depth, seed=3 * 100000 + n, guiding=C.pointer(go)),
                            stats=False)["image"]
            elapsed += time.perf_counter() - t
            acc += img
            n += 1
            curve.append((elapsed * 1e3, scenes.rel_mse(acc / n, ref)))
            if curve[-1][1] <= threshold * 0.8:
                break
        to = _time_to(curve, threshold)
        # relMSE at the plain run's spp count, and the cost of a guided sample
        return {"guided_ms_to_threshold": None if to is None else round(to, 1),
                "speedup": None if (tp is None or to is None) else round(tp / to, 3), "guided_spp_run": n * chunk,
                "ms_per_spp_vs_plain": round((elapsed / (n * chunk)) / (pc[-1][0] * 1e-3 / pspp), 3)}

    ceiling = {"fit": "offline: 200 training renders of 8 spp, 4 Adam steps of 2^18 records each "
                      "(lr 1e-2, 3e-3, 1e-3), then frozen; time excludes the fit"}
    if full and analysis:
        for gd in (1, 0):
            for alpha in (0.5, 0.75):
                ceiling[f"guide_depth_{gd}_alpha_{alpha}"] = offline(gd, alpha)
        print(f"# guided leg: offline-guide ceiling {ceiling}", file=sys.stderr, flush=True)
    return {
        "workload": f"narrow-aperture room, displaced icosphere level 8, {res}x{res}, depth 8, "
                    f"beam through a 0.3x0.3 gap onto the ceiling",
        "judge_spp": judge_spp, "threshold": threshold,
        "plain_ms_to_threshold": None if tp is None else round(tp, 1),
        "guided_ms_to_threshold": None if tg is None else round(tg, 1),
        "speedup": None if (tp is None or tg is None) else round(tp / tg, 3),
        "plain_spp_run": pspp, "guided_spp_run": gspp,
        "guided_training": f"online, first {train_calls} calls of {chunk} spp, 4 Adam steps of 2^18 per pass",
        "guided_vertices": "all" if guide_depth == 0 else f"the first {guide_depth}",
        "training_targets": "f |cos| L_i / pdf" if product_targets else "L_i / pdf",
        "every_vertex_guided": {"guided_ms_to_threshold": None if ta is None else round(ta, 1),
                                "speedup": None if (tp is None or ta is None) else round(tp / ta, 3),
                                "guided_spp_run": aspp},
        "offline_guide_ceiling": ceiling if full and analysis else None,
        "variance_by_depth": variance_by_depth(tr, product_targets) if full and analysis else None,
    }


def variance_by_depth(tr, product_targets, res=256, judge_spp=4096, spp=256):
    """Where the narrow-aperture room's variance lives: for max depth D, the per-sample
    variance of the plain estimator (mean over pixels of the squared error at `spp` times
    spp, less the judge's noise; judges of 2 x judge_spp spp at the same depth), and the
    ratio of a guided estimator's to it, the guide fitted offline at that depth (200
    training renders of 8 spp) and frozen, at the first vertex or at every vertex, mixing
    fraction 0.5 and 0.9.  Depth 1 is direct light only (next-event estimation), depth 2 adds
    the first indirect bounce, which a guide at the first vertex steers."""
    import numpy as np
    from rainbow_amd import scenes
    from rainbow_amd.tracer import guiding_opts, render_params
    sc, cam = scenes.narrow_aperture_room(level=8, width=res, height=res, light_tilt=45.0)
    tr.upload(sc)
    out = {"workload": f"narrow-aperture room, level 8, {res}x{res}", "judge_spp": 2 * judge_spp, "spp": spp}
    for depth in (1, 2, 3, 8):
        r1 = tr.render(cam, render_params(spp=judge_spp, max_depth=depth, seed=1 << 20), stats=False)["image"]
        r2 = tr.render(cam, render_params(spp=judge_spp, max_depth=depth, seed=1 << 21), stats=False)["image"]
        r1, r2 = r1.astype(np.float64), r2.astype(np.float64)
        ref = 0.5 * (r1 + r2)
        noise = float(np.mean((r1 - r2) ** 2)) / 4

        def var(img):
            return (float(np.mean((img - ref) ** 2)) - noise) * spp

        vp = var(tr.render(cam, render_params(spp=spp, max_depth=depth, seed=77), stats=False)["image"])
        row = {"mean": round(float(ref.mean()), 5), "plain_variance": float(f"{vp:.4g}")}
        for gd in (1, 0):
            tr.guide_reset(7)
            for n in range(200):
                lr = 1e-2 if n < 100 else (3e-3 if n < 150 else 1e-3)
                go = guiding_opts(train=True, steps_per_pass=4, batch=1 << 18, guide_depth=gd, learning_rate=lr,
                                  product_targets=product_targets)
                tr.render(cam, render_params(spp=8, max_depth=depth, seed=5000 + n, guiding=C.pointer(go)), stats=False)
            for alpha in (0.5, 0.9):
                go = guiding_opts(mis_fraction=alpha, train=False, guide_depth=gd)
                img = tr.render(cam, render_params(spp=spp, max_depth=depth, seed=99, guiding=C.pointer(go)),
                                stats=False)["image"]
                row[f"guide_depth_{gd}_alpha_{alpha}_ratio"] = round(var(img) / vp, 3)
        out[f"depth_{depth}"] = row
    print(f"# guided leg: variance by depth {out}", file=sys.stderr, flush=True)
    return out


def volume_leg(cpu=True, steps=2):
    """The volume path integrator on the foggy Cornell box (a homogeneous fog sphere and a
    16^3 smoke grid behind invisible boundaries), 512x512, 64 spp, depth 8, beside the
    oracle's rate on a bounded sample of the same workload."""
    from rainbow_amd import abi, scenes
    from rainbow_amd.tracer import Tracer, render_params
    sc, cam = scenes.foggy_cornell_box(512, 512, grid=True)
    spp, depth = 64, 8
    tr = Tracer(0)
    tr.upload(sc)
    params = render_params(spp=spp, max_depth=depth, seed=0, integrator=abi.RBW_INTEGRATOR_VOLUME_PATH)
    tr.render(cam, params)
    rays, samples = 0, 0
    t0 = time.perf_counter()
    for _ in range(steps):
        st_ = abi.rbw_stats()
        rc = tr.lib.rbw_render(tr.h, C.byref(cam), C.byref(params), None, C.byref(st_))
        if rc != 0:
            raise RuntimeError(tr.lib.rbw_last_error().decode())
        d = st_.as_dict()
        samples += d["samples"]
        rays += d["extension_rays"] + d["shadow_rays"] + d["mis_rays"]
    wall = time.perf_counter() - t0
    out = {"workload": "Cornell box with a homogeneous fog sphere (g 0.5) and a 16^3 smoke grid (g -0.2), "
                       "512x512, 64 spp, max depth 8, reference MIS weighting",
           "value": round(samples / wall / 1e6, 3), "unit": "Msamples/s",
           "mrays_per_s": round(rays / wall / 1e6, 2), "ms_per_render": round(wall / steps * 1e3, 2)}
    print(f"# volume leg: {out['value']} Msamples/s", file=sys.stderr, flush=True)
    if cpu:
        cb = cpu_baseline(sc, cam, spp, depth, target_s=5.0, integrator=abi.RBW_INTEGRATOR_VOLUME_PATH)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = round(out["value"] / cb["value"], 1)
    return out


def mlp_leg(rows=1 << 18, calls=20):
    """The guiding network's kernels at the training batch of 2^18 records: one training
    step (forward with stored activations, mixture gradient, data and weight gradients,
    Adam) and batched inference, each timed with HIP events on the library's stream
    (transfers excluded), against the dense FP16 MFMA peak.  Records are synthetic:
    inputs uniform in [0,1]^9, directions uniform on the sphere, weights exponential."""
    import numpy as np

    from rainbow_amd import roofline
    from rainbow_amd.tracer import Tracer
    rng = np.random.default_rng(7)
    x = rng.random((rows, 9), dtype=np.float32)
    d = rng.normal(size=(rows, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    c = rng.exponential(size=rows).astype(np.float32)
    tr = Tracer(0)
    tr.guide_reset(1)
    for _ in range(3):
        tr.guide_train(x, d, c, want_loss=False)
        tr.guide_forward(x)
    train_ms, fwd_ms = [], []
    for _ in range(calls):
        tr.guide_train(x, d, c, want_loss=False)
        train_ms.append(tr.guide_last_ms())
    for _ in range(calls):
        tr.guide_forward(x)
        fwd_ms.append(tr.guide_last_ms())
    tr.close()
    out = {"workload": f"{rows} synthetic records, 9 inputs -> 216 sin/cos features -> 5 x 64 ReLU -> 8 vMF lobes",
           "peak": roofline.MFMA_FP16_PEAK_TFLOPS, "unit": "TFLOP/s", "bound": "mfma"}
    for name, ms, flop in (("train_step", train_ms, roofline.MLP_FLOP_TRAIN),
                           ("inference", fwd_ms, roofline.MLP_FLOP_FORWARD)):
        m = float(np.median(ms))
        tf = rows * flop / (m * 1e-3) / 1e12
        out[name] = {"ms": round(m, 4), "flop_per_record": flop, "achieved": round(tf, 1),
                     "frac": round(tf / roofline.MFMA_FP16_PEAK_TFLOPS, 4)}
    print(f"# mlp leg: train step {out['train_step']['ms']} ms, inference {out['inference']['ms']} ms",
          file=sys.stderr, flush=True)
    return out


def integrator_variants_leg(sc, cam, spp, depth, cpu=True, steps=2):
    """The headline workload rendered by the other estimators the build carries: the direct
    integrator (one light and one BSDF sample, beside the oracle's rate) and the path
    integrator with the stratified sampler (8 x 8 strata, 8 1D and 12 2D dimensions)."""
    from rainbow_amd import abi
    from rainbow_amd.tracer import render_params

    def rate(tr, params):
        tr.render(cam, params, stats=False)
        samples = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            st_ = abi.rbw_stats()
            rc = tr.lib.rbw_render(tr.h, C.byref(cam), C.byref(params), None, C.byref(st_))
            if rc != 0:
                raise RuntimeError(tr.lib.rbw_last_error().decode())
            samples += st_.as_dict()["samples"]
        return samples / (time.perf_counter() - t0) / 1e6

    from rainbow_amd.tracer import Tracer
    tr = Tracer(0)
    tr.upload(sc)
    direct = render_params(spp=spp, max_depth=depth, seed=0, integrator=abi.RBW_INTEGRATOR_DIRECT,
                           emitter_samples=1, bsdf_samples=1)
    strat = render_params(spp=spp, max_depth=depth, seed=0, sampler=abi.RBW_SAMPLER_STRATIFIED, strata_x=8,
                          strata_y=spp // 8, dims_1d=8, dims_2d=12)
    out = {"direct": {"value": round(rate(tr, direct), 2), "unit": "Msamples/s",
                      "workload": "the headline scene, direct integrator, 1 light + 1 BSDF sample"},
           "stratified": {"value": round(rate(tr, strat), 2), "unit": "Msamples/s",
                          "workload": f"the headline workload, stratified sampler 8x{spp // 8}, 8 1D + 12 2D dims"}}
    print(f"# direct {out['direct']['value']} Msamples/s, stratified {out['stratified']['value']} Msamples/s",
          file=sys.stderr, flush=True)
    if cpu:
        cb = cpu_baseline(sc, cam, spp, depth, target_s=5.0, integrator=abi.RBW_INTEGRATOR_DIRECT,
                          emitter_samples=1, bsdf_samples=1)
        out["direct"]["cpu_baseline"] = cb
        out["direct"]["gpu_over_cpu"] = round(out["direct"]["value"] / cb["value"], 1)
    return out


def rooms_batch_leg(n=16, res=(256, 1024), judge_spp=1024, threshold=1e-3, chunk=8, max_spp=4096, train_calls=32,
                    guide_depth=1, product_targets=True, pool=0, steps_per_pass=1):
    """The guided-plus-batched workload: 16 narrow-aperture room variants (gap 0.1-0.5,
    1e5-1.3e6 triangles, GGX alpha 0.02-0.6, 256^2-1024^2) submitted as ONE batch of jobs,
    guided (one network per job, all trained online by one batched kernel, the first
    guide_depth vertices guided) against the same batch with the plain estimator.  A call
    renders `chunk` spp of every job; time is the calls' wall time (tracing, inference,
    training).  A job's relMSE is measured against a judge of 2 x judge_spp spp (two
    independent halves R1, R2), less the judge's own noise, mean((R1 - R2)^2 / (R^2 + 0.01)) / 4,
    so a modest judge does not bias the estimate.  Reports each job's time to relMSE <=
    threshold and the time until every job is below it.  pool: path slots of both schedules
    (0: the library's default); steps_per_pass: Adam steps per training point (every 8 bounce
    iterations)."""
    import numpy as np
    from rainbow_amd import scenes
    from rainbow_amd.tracer import Tracer, batch_opts, guiding_opts
    t0 = time.perf_counter()
    rooms = scenes.room_variants(n=n, res=res)
    tr = Tracer(0)
    ids = [tr.add_scene(r["scene"]) for r in rooms]
    print(f"# rooms batch: {n} variants built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    depth = 8

    def jobs(spp, seed0):
        return [{"scene": ids[k], "camera": r["camera"], "spp": spp, "max_depth": depth, "seed": seed0 + 131 * k}
                for k, r in enumerate(rooms)]

    t0 = time.perf_counter()
    r1, _, _ = tr.submit_jobs(jobs(judge_spp, 1 << 20), batch_opts())
    r2, _, _ = tr.submit_jobs(jobs(judge_spp, 1 << 21), batch_opts())
    refs = [(a + b) * 0.5 for a, b in zip(r1, r2)]
    noise = [float(np.mean((a.astype(np.float64) - b) ** 2 / (r.astype(np.float64) ** 2 + 0.01))) / 4
             for a, b, r in zip(r1, r2, refs)]
    print(f"# rooms batch: judges ({2 * judge_spp} spp) in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)

    def run(guided):
        acc = [np.zeros_like(r) for r in refs]
        curves = [[] for _ in refs]
        elapsed, calls, steps = 0.0, 0, 0
        if guided:
            tr.guide_jobs_reset(n, seed=7)
        while calls * chunk < max_spp:
            o = batch_opts(pool_size=pool)
            if guided:
                o = batch_opts(pool_size=pool,
                               guiding=guiding_opts(train=calls < train_calls, steps_per_pass=steps_per_pass,
                                                    batch=1 << 16, guide_depth=guide_depth,
                                                    product_targets=product_targets))
            t = time.perf_counter()
            imgs, _, st = tr.submit_jobs(jobs(chunk, (2 if guided else 1) * 100000 + 1000 * calls), o)
            elapsed += time.perf_counter() - t
            steps += st["guide_steps"]
            calls += 1
            worst = 0.0
            for k, img in enumerate(imgs):
                acc[k] += img
                e = scenes.rel_mse(acc[k] / calls, refs[k]) - noise[k]
                curves[k].append((elapsed * 1e3, max(e, 1e-12)))
                worst = max(worst, e)
            if calls % 64 == 0:
                print(f"# rooms batch: {'guided' if guided else 'plain'} {calls * chunk} spp, {elapsed:.1f} s, "
                      f"worst relMSE {worst:.3g}", file=sys.stderr, flush=True)
            if worst <= threshold * 0.8:
                break
        return curves, calls * chunk, steps

    pc, pspp, _ = run(False)
    gc, gspp, gsteps = run(True)
    per, tp_all, tg_all = [], 0.0, 0.0
    for k, r in enumerate(rooms):
        tp, tg = _time_to(pc[k], threshold), _time_to(gc[k], threshold)
        t_eq = min(pc[k][-1][0], gc[k][-1][0])
        ep, eg = _err_at(pc[k], t_eq), _err_at(gc[k], t_eq)
        per.append({"gap": round(r["gap"], 3), "triangles": r["triangles"], "res": r["res"],
                    "plain_ms": None if tp is None else round(tp, 1), "guided_ms": None if tg is None else round(tg, 1),
                    "speedup": None if (tp is None or tg is None) else round(tp / tg, 3),
                    "equal_time_ms": round(t_eq, 1), "plain_relmse": float(f"{ep:.4g}"),
                    "guided_relmse": float(f"{eg:.4g}"), "relmse_ratio": round(eg / ep, 3)})
        tp_all = None if tp is None or tp_all is None else max(tp_all, tp)
        tg_all = None if tg is None or tg_all is None else max(tg_all, tg)
    sp = sorted(p["speedup"] for p in per if p["speedup"])
    ratios = sorted(p["relmse_ratio"] for p in per)
    out = {"workload": f"{n} narrow-aperture room variants (gap 0.1-0.5, 1e5-1.3e6 tris, GGX alpha 0.02-0.6, "
                       f"{res[0]}^2-{res[1]}^2) in one submission, depth {depth}, one network per job",
           "judge": f"2 x {judge_spp} spp per job, relMSE less the judge's noise", "threshold": threshold,
           "guided_vertices": "all" if guide_depth == 0 else f"the first {guide_depth}",
           "training": f"online in the first {train_calls} calls of {chunk} spp: {steps_per_pass} batched Adam "
                       f"step(s) (2^16 rows per network) after every 8 bounce iterations ({gsteps} steps), targets "
                       + ("f |cos| L_i / pdf" if product_targets else "L_i / pdf"),
           "pool_size": pool or "default",
           "plain_ms_all_jobs": None if tp_all is None else round(tp_all, 1),
           "guided_ms_all_jobs": None if tg_all is None else round(tg_all, 1),
           "speedup_all_jobs": None if (tp_all is None or tg_all is None) else round(tp_all / tg_all, 3),
           "median_speedup": sp[len(sp) // 2] if sp else None, "min_speedup": sp[0] if sp else None,
           "max_speedup": sp[-1] if sp else None, "jobs_crossing": {"plain": sum(p["plain_ms"] is not None for p in per),
                                                                     "guided": sum(p["guided_ms"] is not None for p in per)},
           "median_relmse_ratio_equal_time": ratios[len(ratios) // 2], "plain_spp_run": pspp, "guided_spp_run": gspp, "per_job": per}
    print(f"# rooms batch: all jobs plain {out['plain_ms_all_jobs']} ms, guided {out['guided_ms_all_jobs']} ms",
          file=sys.stderr, flush=True)
    for i in ids:
        tr.remove_scene(i)
    tr.close()
    return out


def batch_leg(n_jobs=64, autotune=True):
    """The batched-jobs workload: n_jobs seeded jobs of uneven size (grid-sphere meshes of
    1e4-2e6 triangles log-uniform in the Cornell box, 128^2-1024^2, 4-256 spp, depth 3-12)
    submitted as one batch.  Reports aggregate and per-job Msamples/s and, for comparison,
    the same jobs rendered one rbw_render call at a time."""
    from rainbow_amd import scenes
    from rainbow_amd.tracer import Tracer, batch_opts, render_params
    t0 = time.perf_counter()
    jobs = scenes.batch_workload(n_jobs=n_jobs, seed=2)
    tr = Tracer(0)
    ids = [tr.add_scene(j["scene"]) for j in jobs]
    build_s = time.perf_counter() - t0
    print(f"# batch leg: {n_jobs} jobs generated and BVH-built in {build_s:.1f} s", file=sys.stderr, flush=True)
    js = [{"scene": ids[i], "camera": j["camera"], "spp": j["spp"], "max_depth": j["max_depth"], "seed": j["seed"]}
          for i, j in enumerate(jobs)]
    opts = batch_opts(autotune=autotune)
    tuned_ms = 0.0
    if autotune:
        opts, tuned_ms = tr.autotune_jobs(js, opts)
        opts.autotune = 0
    tr.submit_jobs(js, opts, images=False)  # warm: merged scene upload, pool allocation
    _, jst, bst = tr.submit_jobs(js, opts, images=False)
    print(f"# batch: {bst['msamples_per_s']:.1f} Msamples/s in {bst['render_ms']:.0f} ms, "
          f"{bst['iterations']} iterations", file=sys.stderr, flush=True)
    # the same jobs one render call at a time (scene upload excluded)
    single_ms = 0.0
    for j in jobs:
        tr.upload(j["scene"])
        st = tr.render(j["camera"], render_params(spp=j["spp"], max_depth=j["max_depth"], seed=j["seed"]),
                       stats=True)["stats"]
        single_ms += st["render_ms"]
    samples = bst["samples"]
    per_job = [round(x["msamples_per_s"], 1) for x in jst]
    # the jobs that share the persistent launch (fewer samples than the pool has slots):
    # as one submission against one render call each
    pool = bst["pool_size"]
    small = [k for k, j in enumerate(jobs)
             if j["camera"].width * j["camera"].height * j["spp"] < pool]
    small_out = None
    if len(small) > 1:
        sj = [js[k] for k in small]
        tr.submit_jobs(sj, opts, images=False)
        _, _, sb = tr.submit_jobs(sj, opts, images=False)
        one_ms = 0.0
        for k in small:
            j = jobs[k]
            tr.upload(j["scene"])
            one_ms += tr.render(j["camera"], render_params(spp=j["spp"], max_depth=j["max_depth"], seed=j["seed"]),
                                stats=True)["stats"]["render_ms"]
        small_out = {"jobs": len(small), "samples": sb["samples"], "batch_ms": round(sb["render_ms"], 1),
                     "one_at_a_time_ms": round(one_ms, 1), "gain": round(one_ms / sb["render_ms"], 3)}
        print(f"# batch: {len(small)} pool-sharing jobs: batch {sb['render_ms']:.1f} ms, one at a time "
              f"{one_ms:.1f} ms", file=sys.stderr, flush=True)
    return {
        "workload": f"{n_jobs} jobs, seed 2: grid-sphere mesh 1e4-2e6 tris (log-uniform) in the Cornell box, "
                    "128^2-1024^2, 4-256 spp, depth 3-12, per-triangle Lambert/GGX",
        "jobs": n_jobs, "samples": samples, "triangles": sum(j["triangles"] for j in jobs),
        "batch_ms": round(bst["render_ms"], 1), "aggregate_msamples_per_s": round(bst["msamples_per_s"], 2),
        "mrays_per_s": round(bst["mrays_per_s"], 1),
        "one_at_a_time_ms": round(single_ms, 1),
        "one_at_a_time_msamples_per_s": round(samples / single_ms * 1e-3, 2),
        "per_job_msamples_per_s": {"min": min(per_job), "median": sorted(per_job)[len(per_job) // 2],
                                   "max": max(per_job)},
        "iterations": bst["iterations"], "graph_launches": bst["graph_launches"],
        "block_size": bst["block_size"], "pool_size": bst["pool_size"], "bvh_width": bst["bvh_width"],
        "autotune_ms": round(tuned_ms, 1),
        "tuned": {"pool_size": opts.pool_size, "block_size": opts.block_size, "bvh_width": opts.bvh_width},
        "pool_sharing_jobs": small_out,
    }


PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                       "tracer_persistent_bvh4_icosphere_1024_64spp.md")


def profiled_kernels(stage, width):
    """the profile's names of the persistent-slot kernels of a stage (extend, shade, connect)
    at a BVH width on the headline scene (one planar light): at width 4 the extension walk
    runs persistent, shading (grouped by material kind) also tests the MIS rays against the
    emitter triangles, and the connect stage is the persistent occlusion walk and the
    walk-free connect kernel; at widths 2 and 8 connect is the occlusion-only instantiation"""
    if stage == "extend":
        return ("k_b_extend_dyn",) if width == 4 else (f"k_b_extend<{width}>",)
    if stage == "connect":
        if width == 4:
            return ("k_b_occlusion_dyn", "k_b_connect<4, true, true>")
        return (f"k_b_connect<{width}, true, false>",)
    return ("k_b_shade_sorted<true>",) if width == 4 else ("k_b_shade_sorted<false>",)


def profiled_traffic(kernels):
    """HBM bytes per launch of the stage made of `kernels` on the icosphere workload, from
    the committed rocprofv3 summary (FETCH_SIZE and WRITE_SIZE runs, each scaled by what the
    counter reads for a copy kernel of known size); None when the summary lacks a row."""
    try:
        lines = open(PROFILE).read().splitlines()
    except OSError:
        return None
    rows = {}
    for line in lines:
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0] in kernels:
            try:
                rows[cells[0]] = float(cells[4]) * 1e6
            except ValueError:
                return None
    if len(rows) != len(kernels):
        return None
    return int(sum(rows.values()))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="icosphere", choices=["icosphere", "cornell"])
    ap.add_argument("--bvh-width", type=int, default=0, choices=[0, 2, 4, 8],
                    help="BVH width of the headline renders (0: the library's default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-counters", action="store_true", help="skip the instrumented (visit-count) run")
    ap.add_argument("--batch", action=argparse.BooleanOptionalAction, default=True,
                    help="batched-jobs leg (64 jobs of uneven size in one submission)")
    ap.add_argument("--variants", action=argparse.BooleanOptionalAction, default=True,
                    help="direct integrator and stratified sampler on the headline workload")
    ap.add_argument("--rooms", action=argparse.BooleanOptionalAction, default=True,
                    help="guided vs plain on 16 narrow-aperture room variants in one batch, one network per job")
    ap.add_argument("--volume", action=argparse.BooleanOptionalAction, default=True,
                    help="volume path integrator leg (foggy Cornell box)")
    ap.add_argument("--guide-depth", type=int, default=1,
                    help="guided leg: guide the first N vertices of a path, the plain estimator after (0: all)")
    ap.add_argument("--mlp", action=argparse.BooleanOptionalAction, default=True,
                    help="guiding-network kernels (training step, inference) at 2^18 records vs the MFMA peak")
    ap.add_argument("--product-targets", action=argparse.BooleanOptionalAction, default=True,
                    help="guided legs: train the networks on f |cos| L_i (else L_i)")
    ap.add_argument("--guided-analysis", action="store_true",
                    help="guided leg: also the offline-fitted guide's ceiling and the variance by path depth")
    ap.add_argument("--guided", default="full", choices=["reduced", "full", "off"],
                    help="guided-vs-plain time-to-relMSE leg: 512^2 with a 65536-spp judge (full, the "
                         "configured workload, ~90 s), 256^2 with a 16384-spp judge (reduced), or skipped")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(the film reduction across GPUs is designed but not built)", file=sys.stderr)
        sys.exit(2)
    if not args.no_cpu:
        # the CPU baseline's -O3 -march=native build for this host, compiled by a child process
        # while the GPU legs run (started before anything touches the GPU)
        from oracle import orc_py
        if os.path.exists(orc_py.NATIVE_LIB):
            os.remove(orc_py.NATIVE_LIB)
        global _NATIVE_BUILD
        _NATIVE_BUILD = orc_py.start_native_build()

    from rainbow_amd import abi, roofline, scenes
    from rainbow_amd.tracer import Tracer, render_params

    if args.workload == "icosphere":
        sc, cam = scenes.displaced_icosphere_box(level=8, width=1024, height=1024)
        spp, depth = 64, 8
        workload = ("fBm-displaced icosphere level 8 (1,310,720 tris, per-triangle Lambert/GGX) in the "
                    "Cornell box, 1024x1024, 64 spp, max depth 8, RR from depth 3")
    else:
        sc, cam = scenes.cornell_box()
        spp, depth = 16, 5
        workload = "Cornell box (10 wall tris, 2 spheres, quad light), 256x256, 16 spp, max depth 5"
    ntris = sum(len(i) for i in sc.indices)
    tr = Tracer(0)
    tr.upload(sc)
    params = render_params(spp=spp, max_depth=depth, rr_threshold=1.0, seed=0, bvh_width=args.bvh_width)

    counted = None
    if not args.no_counters:
        cp = render_params(spp=spp, max_depth=depth, rr_threshold=1.0, seed=0, count_visits=1,
                           bvh_width=args.bvh_width)
        counted = tr.render(cam, cp)["stats"]
        print(f"# instrumented run: {counted['node_visits']} node visits, {counted['prim_tests']} prim tests",
              file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        tr.render(cam, params)
    step_stats = []
    import numpy as np
    film = np.zeros((cam.height, cam.width, 3), np.float32)
    t0 = time.perf_counter()
    for i in range(args.steps):
        # one render call per step, delivering the film's pixels to host memory as the
        # drop-in render call does (the PCIe copy is inside the timed region)
        st_ = abi.rbw_stats()
        rc = tr.lib.rbw_render(tr.h, C.byref(cam), C.byref(params), abi.ptr(film, abi.f32p), C.byref(st_))
        if rc != 0:
            raise RuntimeError(tr.lib.rbw_last_error().decode())
        step_stats.append(st_.as_dict())
        print(f"# step {i}: {st_.render_ms:.1f} ms, {st_.msamples_per_s:.1f} Msamples/s", file=sys.stderr,
              flush=True)
    t1 = time.perf_counter()
    assert np.isfinite(film).all() and film.any()
    wall = t1 - t0
    samples = sum(s["samples"] for s in step_stats)
    rays = sum(s["extension_rays"] + s["shadow_rays"] + s["mis_rays"] for s in step_stats)
    value = samples / wall / 1e6
    ms_per_step = wall / args.steps * 1e3

    roof = None
    if counted:
        # per-kernel HIP-event times: one more render of the same workload with the bounce
        # loop launched kernel by kernel (the timed steps replay it as a hipGraph)
        tp = render_params(spp=spp, max_depth=depth, rr_threshold=1.0, seed=0, time_kernels=1,
                           bvh_width=args.bvh_width)
        timed = tr.render(cam, tp)["stats"]
        width = counted["bvh_width"]
        kb = roofline.kernel_bytes(counted, width)
        last = timed
        kms = {k: last["kernel_ms"][k] for k in ("extend", "shade", "connect")}
        dom = max(kms, key=kms.get)
        launches = last["kernel_launches"][dom]
        achieved = kb[dom] / (kms[dom] * 1e-3) / 1e9
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": roofline.HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / roofline.HBM_PEAK_GBS, 4),
                "traffic": profiled_traffic(profiled_kernels(dom, width)) if args.workload == "icosphere" else None,
                "bvh_width": width,
                "algorithmic_bytes_per_launch": int(kb[dom] / max(1, launches)),
                "avg_launch_ms": kms[dom] / max(1, launches),
                "kernel_ms": {k: round(v, 3) for k, v in kms.items()},
                # every stage kernel against the same roofline (the dominant one is above)
                "kernels": {k: {"achieved": round(kb[k] / (kms[k] * 1e-3) / 1e9, 1),
                                "frac": round(kb[k] / (kms[k] * 1e-3) / 1e9 / roofline.HBM_PEAK_GBS, 4),
                                "avg_launch_ms": kms[k] / max(1, last["kernel_launches"][k])} for k in kms}}

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "triangles": ntris, "width": cam.width, "height": cam.height,
                   "spp": spp, "max_depth": depth, "sampler": "per-sample pcg32", "triangle_test": "moller-trumbore",
                   "shadow_rays": "closest-visible results (any-hit walk: the light is planar)",
                   "mis_rays": "closest-visible results (emitter triangles first, then an occlusion walk)"},
        "mrays_per_s": round(rays / wall / 1e6, 2),
        # the render on the device alone (HIP events: first kernel to film resolve), without
        # the film's copy to host memory that ms_per_step includes
        "render_ms_device": round(sum(s["render_ms"] for s in step_stats) / len(step_stats), 3),
        "roofline": roof,
    }
    if args.batch:
        out["batch"] = batch_leg()
    if args.guided != "off":
        out["guided"] = guided_leg(args.guided == "full", guide_depth=args.guide_depth,
                                   product_targets=args.product_targets, analysis=args.guided_analysis)
    if args.volume:
        out["volume"] = volume_leg(cpu=not args.no_cpu)
    if args.mlp:
        out["guiding_mlp"] = mlp_leg()
    if args.variants:
        out["integrator_variants"] = integrator_variants_leg(sc, cam, spp, depth, cpu=not args.no_cpu)
    if args.rooms:
        out["room_variants"] = rooms_batch_leg(guide_depth=args.guide_depth, product_targets=args.product_targets)
    if not args.no_cpu:
        cb = cpu_baseline(sc, cam, spp, depth)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = round(value / cb["value"], 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
