"""Python handle over librbw.so: the MI355X wavefront path tracer.

Mirrors the reference's render call (integrator::render, rainbow-cpu/
integrators/integrator.hpp:64-66): a scene is uploaded once, then render()
fills a film for a camera.  Every call goes through the C ABI; there is no
Python or CPU fallback -- a missing library or device is an error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi


class RbwError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"rbw error {code}: {msg}")
        self.code = code


_lib = None


def library():
    global _lib
    if _lib is None:
        _lib = abi.load_library()
    return _lib


def _check(rc):
    if rc != abi.RBW_OK:
        raise RbwError(rc, library().rbw_last_error().decode())


class Tracer:
    def __init__(self, device=0):
        self.lib = library()
        self.h = C.c_void_p()
        opts = abi.rbw_device_opts(device=device)
        _check(self.lib.rbw_create(C.byref(opts), C.byref(self.h)))
        self._scene_desc = None

    def close(self):
        if self.h:
            self.lib.rbw_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, scene):
        desc = scene.desc()
        _check(self.lib.rbw_upload_scene(self.h, C.byref(desc)))
        self._scene_desc = desc
        return self

    def upload_desc(self, desc):
        """uploads a caller-built rbw_scene_desc"""
        _check(self.lib.rbw_upload_scene(self.h, C.byref(desc)))

    def render(self, cam, params, debug=False, stats=True):
        W, H = cam.width, cam.height
        out = np.zeros((H, W, 3), np.float32)
        dbg = None
        if debug:
            dbg = np.full(W * H * params.spp * params.max_depth * 3, abi.RBW_NONE, np.uint32)
            params.debug_hits = abi.ptr(dbg, abi.u32p)
        st = abi.rbw_stats()
        try:
            _check(self.lib.rbw_render(self.h, C.byref(cam), C.byref(params), abi.ptr(out, abi.f32p),
                                       C.byref(st) if stats else None))
        finally:
            if debug:
                params.debug_hits = abi.ptr(None, abi.u32p)
        res = {"image": out, "stats": st.as_dict()}
        if debug:
            res["debug_hits"] = dbg.reshape(H * W * params.spp, params.max_depth, 3)
        return res

    def trace(self, origins, directions, tmax, visible_only=False, triangle_test=abi.RBW_TRI_MOLLER_TRUMBORE,
              bvh_width=0, static=False):
        """closest-hit / closest-visible queries; static: through rbw_trace_rays_static (no
        device allocation, at most abi.RBW_STATIC_MAX_RAYS rays)"""
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(directions, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(tmax, np.float32).reshape(-1)
        out = np.zeros(len(o), abi.HIT_DTYPE)
        fn = self.lib.rbw_trace_rays_static if static else self.lib.rbw_trace_rays
        _check(fn(self.h, len(o), abi.ptr(o, abi.f32p), abi.ptr(d, abi.f32p), abi.ptr(t, abi.f32p), int(visible_only),
                  int(triangle_test), int(bvh_width), out.ctypes.data_as(C.c_void_p)))
        return out


    # -- guiding network (no counterpart in the reference) -------------------
    def sampler_draws(self, params, pixel, n1, n2):
        """per sample of `pixel`: n1 1D draws, then n2 2D draws -> (spp, n1 + 2*n2)"""
        out = np.zeros((params.spp, n1 + 2 * n2), np.float32)
        _check(self.lib.rbw_sampler_draws(self.h, C.byref(params), pixel, n1, n2, abi.ptr(out, abi.f32p)))
        return out

    def guide_reset(self, seed=0):
        _check(self.lib.rbw_guide_reset(self.h, seed))

    def guide_params(self):
        out = np.zeros(abi.RBW_GUIDE_PARAMS, np.float32)
        _check(self.lib.rbw_guide_get_params(self.h, abi.ptr(out, abi.f32p)))
        return out

    def guide_set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        assert flat.size == abi.RBW_GUIDE_PARAMS
        _check(self.lib.rbw_guide_set_params(self.h, abi.ptr(flat, abi.f32p)))

    def guide_forward(self, inputs):
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 9)
        out = np.zeros((len(x), 48), np.float32)
        _check(self.lib.rbw_guide_forward(self.h, len(x), abi.ptr(x, abi.f32p), abi.ptr(out, abi.f32p)))
        return out

    def guide_train(self, inputs, dirs, c, lr=1e-2, apply=True, want_grad=False, want_loss=True):
        x = np.ascontiguousarray(inputs, np.float32).reshape(-1, 9)
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        w = np.ascontiguousarray(c, np.float32).reshape(-1)
        loss = np.zeros(1, np.float32) if want_loss else None
        grad = np.zeros(abi.RBW_GUIDE_PARAMS, np.float32) if want_grad else None
        _check(self.lib.rbw_guide_train(self.h, len(x), abi.ptr(x, abi.f32p), abi.ptr(d, abi.f32p),
                                        abi.ptr(w, abi.f32p), float(lr), int(apply), abi.ptr(loss, abi.f32p),
                                        abi.ptr(grad, abi.f32p)))
        return (float(loss[0]) if want_loss else None), grad

    def guide_last_ms(self):
        """HIP-event time of the network kernels of the last guide_forward / guide_train call"""
        ms = C.c_double(0.0)
        _check(self.lib.rbw_guide_last_ms(self.h, C.byref(ms)))
        return ms.value


    # -- batched jobs (no counterpart in the reference) ----------------------
    def add_scene(self, scene):
        """adds a scene to the context's pool; returns its id"""
        desc = scene.desc()
        sid = C.c_uint32()
        _check(self.lib.rbw_add_scene(self.h, C.byref(desc), C.byref(sid)))
        return sid.value

    def remove_scene(self, sid):
        _check(self.lib.rbw_remove_scene(self.h, sid))

    # -- the job networks of guided batches (job i: network i) ---------------
    def guide_jobs_reset(self, n_networks, seed=0):
        _check(self.lib.rbw_guide_jobs_reset(self.h, n_networks, seed))

    def guide_jobs_params(self, network):
        out = np.zeros(abi.RBW_GUIDE_PARAMS, np.float32)
        _check(self.lib.rbw_guide_jobs_get_params(self.h, network, abi.ptr(out, abi.f32p)))
        return out

    def guide_jobs_set_params(self, network, params):
        p = np.ascontiguousarray(params, np.float32)
        _check(self.lib.rbw_guide_jobs_set_params(self.h, network, abi.ptr(p, abi.f32p)))

    def guide_jobs_train(self, batches, lr=1e-2, apply=False):
        """one batched training step: batches[k] = (inputs n_k x 9, dirs n_k x 3, c n_k) of
        network k; returns (losses, gradients n_networks x RBW_GUIDE_PARAMS)"""
        rows = np.array([len(b[2]) for b in batches], np.uint32)
        x = np.ascontiguousarray(np.concatenate([b[0] for b in batches]), np.float32)
        d = np.ascontiguousarray(np.concatenate([b[1] for b in batches]), np.float32)
        c = np.ascontiguousarray(np.concatenate([b[2] for b in batches]), np.float32)
        loss = np.zeros(len(batches), np.float32)
        grad = np.zeros((len(batches), abi.RBW_GUIDE_PARAMS), np.float32)
        _check(self.lib.rbw_guide_jobs_train(self.h, len(batches), abi.ptr(rows, abi.u32p), abi.ptr(x, abi.f32p),
                                             abi.ptr(d, abi.f32p), abi.ptr(c, abi.f32p), lr, int(apply),
                                             abi.ptr(loss, abi.f32p), abi.ptr(grad, abi.f32p)))
        return loss, grad

    def submit_jobs(self, jobs, opts=None, images=True):
        """jobs: list of dicts {scene, camera, spp, max_depth, rr_threshold=1, seed=0}.
        Returns (images or None, per-job stats, batch stats)."""
        n = len(jobs)
        arr = (abi.rbw_job * n)()
        outs = []
        for i, j in enumerate(jobs):
            cam = j["camera"]
            arr[i].scene = j["scene"]
            arr[i].spp = j["spp"]
            arr[i].max_depth = j["max_depth"]
            arr[i].rr_threshold = j.get("rr_threshold", 1.0)
            arr[i].seed = j.get("seed", 0)
            arr[i].camera = cam
            if images:
                img = np.zeros((cam.height, cam.width, 3), np.float32)
                outs.append(img)
                arr[i].out_rgb = abi.ptr(img, abi.f32p)
        js = (abi.rbw_job_stats * n)()
        bs = abi.rbw_batch_stats()
        o = opts if opts is not None else batch_opts()
        _check(self.lib.rbw_submit_jobs(self.h, arr, n, C.byref(o), js, C.byref(bs)))
        return (outs if images else None), [js[i].as_dict() for i in range(n)], bs.as_dict()

    def autotune_jobs(self, jobs, opts=None):
        n = len(jobs)
        arr = (abi.rbw_job * n)()
        for i, j in enumerate(jobs):
            arr[i].scene = j["scene"]
            arr[i].spp = j["spp"]
            arr[i].max_depth = j["max_depth"]
            arr[i].rr_threshold = j.get("rr_threshold", 1.0)
            arr[i].seed = j.get("seed", 0)
            arr[i].camera = j["camera"]
        o = opts if opts is not None else batch_opts()
        ms = C.c_double()
        _check(self.lib.rbw_autotune_jobs(self.h, arr, n, C.byref(o), C.byref(ms)))
        return o, ms.value


def batch_opts(block_size=0, pool_size=0, use_graph=True, autotune=False, guiding=None, **kw):
    """guiding: an rbw_guiding_opts (kept alive by the returned options) for a guided batch"""
    o = abi.rbw_batch_opts(block_size=block_size, pool_size=pool_size, bvh_width=0, use_graph=int(use_graph),
                           autotune=int(autotune))
    if guiding is not None:
        o._guiding_ref = guiding
        o.guiding = C.pointer(guiding)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def guiding_opts(mis_fraction=0.5, train=True, learning_rate=1e-2, batch=1 << 18, steps_per_pass=4,
                 light_records=False, guide_depth=0, product_targets=False):
    return abi.rbw_guiding_opts(enable=1, mis_fraction=mis_fraction, train=int(train), learning_rate=learning_rate,
                                batch=batch, steps_per_pass=steps_per_pass, light_records=int(light_records),
                                guide_depth=guide_depth, product_targets=int(product_targets))


def render_params(spp=16, max_depth=5, rr_threshold=1.0, seed=0, **kw):
    p = abi.rbw_render_params(spp=spp, max_depth=max_depth, rr_threshold=rr_threshold, seed=seed)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _bounds(bound_min, bound_max):
    bmin = None if bound_min is None else (C.c_int32 * 2)(*bound_min)
    bmax = None if bound_max is None else (C.c_int32 * 2)(*bound_max)
    return bmin, bmax


def film_to_rgba8(image, bound_min=None, bound_max=None, scale=1.0):
    """The film's 8-bit RGBA image as film::write makes it (cameras/film.cpp:114-153)."""
    img = np.ascontiguousarray(image, dtype=np.float32)
    h, w = img.shape[:2]
    out = np.empty((h, w, 4), dtype=np.uint8)
    bmin, bmax = _bounds(bound_min, bound_max)
    _check(library().rbw_film_to_rgba8(abi.ptr(img, abi.f32p), w, h, bmin, bmax, scale,
                                       out.ctypes.data_as(C.POINTER(C.c_uint8))))
    return out


def write_png(path, image, bound_min=None, bound_max=None, scale=1.0):
    img = np.ascontiguousarray(image, dtype=np.float32)
    h, w = img.shape[:2]
    bmin, bmax = _bounds(bound_min, bound_max)
    _check(library().rbw_write_png(os.fsencode(path), abi.ptr(img, abi.f32p), w, h, bmin, bmax, scale))


def write_pfm(path, image):
    img = np.ascontiguousarray(image, dtype=np.float32)
    h, w = img.shape[:2]
    _check(library().rbw_write_pfm(os.fsencode(path), abi.ptr(img, abi.f32p), w, h))
