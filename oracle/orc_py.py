"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle (oracle/liborc.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product path (rainbow_amd, librbw.so) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from rainbow_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")
REF_PCG = os.path.join(HERE, "_ref", "libref_pcg.so")

ORC_SAMPLER_TILE, ORC_SAMPLER_PER_SAMPLE, ORC_SAMPLER_REPLAY, ORC_SAMPLER_TILE_STRATIFIED = 0, 1, 2, 3


class orc_render_opts(C.Structure):
    _fields_ = [("sampler", C.c_uint32), ("threads", C.c_uint32), ("tile_stride", C.c_uint32),
                ("record_offsets", abi.u64p), ("record_values", abi.f32p), ("record_capacity", C.c_uint64),
                ("record_count", C.c_uint64), ("strata_x", C.c_uint32), ("strata_y", C.c_uint32),
                ("dims_1d", C.c_uint32), ("dims_2d", C.c_uint32)]


class orc_stats(C.Structure):
    _fields_ = [("seconds", C.c_double), ("samples", C.c_uint64), ("camera_rays", C.c_uint64),
                ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64), ("mis_rays", C.c_uint64),
                ("light_casts", C.c_uint64), ("threads", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liborc.so"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


NATIVE_LIB = os.path.join(HERE, "liborc_native.so")


def start_native_build():
    """starts `make native` (liborc_native.so, -O3 -march=native for this host) as a child
    process and returns it; the CPU baseline's timing build"""
    return subprocess.Popen(["make", "-s", "-C", HERE, "liborc_native.so"], stdout=subprocess.DEVNULL,
                            stderr=subprocess.PIPE)


_lib = None
_native = None


def native_lib():
    """the timing build (None when it was not built); the same C API as lib()"""
    global _native
    if _native is None and os.path.exists(NATIVE_LIB):
        _native = _bind(C.CDLL(NATIVE_LIB))
    return _native


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = _bind(C.CDLL(LIB))
    return _lib


def _bind(L):
    """argument and result types of the oracle's C API on a loaded library"""
    L.orc_scene_create.argtypes = [C.POINTER(abi.rbw_scene_desc), C.POINTER(C.c_void_p)]
    L.orc_scene_destroy.argtypes = [C.c_void_p]
    L.orc_scene_num_bvh_nodes.argtypes = [C.c_void_p]
    L.orc_scene_num_bvh_nodes.restype = C.c_uint32
    L.orc_render.argtypes = [C.c_void_p, C.POINTER(abi.rbw_camera_desc), C.POINTER(abi.rbw_render_params),
                             C.POINTER(orc_render_opts), abi.f32p, C.POINTER(orc_stats)]
    for fn in (L.orc_trace_rays, L.orc_trace_rays_linear):
        fn.argtypes = [C.c_void_p, C.c_uint32, abi.f32p, abi.f32p, abi.f32p, C.c_uint32, C.c_void_p]
    L.orc_camera_rays.argtypes = [C.POINTER(abi.rbw_camera_desc), C.c_uint32, abi.f32p, abi.f32p, abi.f32p,
                                  abi.f32p]
    L.orc_pcg32_reals.argtypes = [C.c_uint64, C.c_uint32, abi.f32p]
    L.orc_pcg32_u32.argtypes = [C.c_uint64, C.c_uint32, abi.u32p]
    L.orc_pcg32_range.argtypes = [C.c_uint64, C.c_uint32, abi.u32p, abi.u32p, abi.u32p]
    L.orc_stratified_tables.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                        abi.f32p, abi.f32p]
    L.orc_stratified_tables.restype = C.c_uint64
    L.orc_concentric_sample_disk.argtypes = [C.c_float, C.c_float, abi.f32p]
    L.orc_cosine_sample_hemisphere.argtypes = [C.c_float, C.c_float, abi.f32p]
    L.orc_uniform_sample_triangle.argtypes = [C.c_float, C.c_float, abi.f32p]
    L.orc_ggx_D.argtypes = [C.c_float, C.c_float, abi.f32p]
    L.orc_ggx_D.restype = C.c_float
    L.orc_ggx_lambda.argtypes = [C.c_float, C.c_float, abi.f32p]
    L.orc_ggx_lambda.restype = C.c_float
    L.orc_ggx_sample_wh.argtypes = [C.c_float, C.c_float, abi.f32p, C.c_float, C.c_float, abi.f32p]
    L.orc_ggx_sample_f.argtypes = [C.c_float, C.c_float, abi.f32p, abi.f32p, abi.f32p, C.c_float, C.c_float,
                                   abi.f32p, abi.f32p, abi.f32p]
    L.orc_ggx_pdf.argtypes = [C.c_float, C.c_float, abi.f32p, abi.f32p]
    L.orc_ggx_pdf.restype = C.c_float
    L.orc_ggx_eval.argtypes = [C.c_float, C.c_float, abi.f32p, abi.f32p, abi.f32p, abi.f32p, abi.f32p]
    L.orc_fresnel_conductor.argtypes = [C.c_float, abi.f32p, abi.f32p, abi.f32p]
    return L


def f3(v):
    return np.ascontiguousarray(np.asarray(v, np.float32).reshape(3))


class OracleScene:
    def __init__(self, scene, desc=None, native=False):
        """native: use the timing build liborc_native.so (must have been built)"""
        self.scene = scene
        self.desc = scene.desc() if desc is None else desc
        self.L = native_lib() if native else lib()
        if self.L is None:
            raise RuntimeError("liborc_native.so is not built (oracle/Makefile target native)")
        self.h = C.c_void_p()
        rc = self.L.orc_scene_create(C.byref(self.desc), C.byref(self.h))
        if rc != 0:
            raise RuntimeError(f"orc_scene_create failed: {rc}")

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "L", None) is not None:
            try:
                self.L.orc_scene_destroy(self.h)
            except Exception:
                pass

    @classmethod
    def from_desc(cls, desc):
        """over a caller-built rbw_scene_desc (its arrays must outlive this call)"""
        return cls(None, desc)

    def num_bvh_nodes(self):
        return self.L.orc_scene_num_bvh_nodes(self.h)

    def trace(self, origins, directions, tmax, visible_only=False, linear=False):
        o = np.ascontiguousarray(origins, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(directions, np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(tmax, np.float32).reshape(-1)
        out = np.zeros(len(o), abi.HIT_DTYPE)
        fn = self.L.orc_trace_rays_linear if linear else self.L.orc_trace_rays
        rc = fn(self.h, len(o), abi.ptr(o, abi.f32p), abi.ptr(d, abi.f32p), abi.ptr(t, abi.f32p),
                int(visible_only), out.ctypes.data_as(C.c_void_p))
        assert rc == 0
        return out

    def render(self, cam, params, sampler=ORC_SAMPLER_TILE, threads=0, tile_stride=1, record=False,
               debug=False, strata=(1, 1), dims=(0, 0)):
        """strata/dims: ORC_SAMPLER_TILE_STRATIFIED's (strata_x, strata_y) and (dims_1d, dims_2d)"""
        W, H = cam.width, cam.height
        out = np.zeros((H, W, 3), np.float32)
        opts = orc_render_opts(sampler=sampler, threads=threads, tile_stride=tile_stride, strata_x=strata[0],
                               strata_y=strata[1], dims_1d=dims[0], dims_2d=dims[1])
        keep = []
        if record:
            n = W * H * params.spp
            offs = np.zeros(n + 1, np.uint64)
            cap = n * (4 + 8 * params.max_depth)
            vals = np.zeros(cap, np.float32)
            opts.record_offsets = abi.ptr(offs, abi.u64p)
            opts.record_values = abi.ptr(vals, abi.f32p)
            opts.record_capacity = cap
            keep += [offs, vals]
        dbg = None
        if debug:
            dbg = np.full(W * H * params.spp * params.max_depth * 3, abi.RBW_NONE, np.uint32)
            params.debug_hits = abi.ptr(dbg, abi.u32p)
        st = orc_stats()
        rc = self.L.orc_render(self.h, C.byref(cam), C.byref(params), C.byref(opts), abi.ptr(out, abi.f32p),
                              C.byref(st))
        if debug:
            params.debug_hits = abi.ptr(None, abi.u32p)
        if rc != 0:
            raise RuntimeError(f"orc_render failed: {rc}")
        res = {"image": out, "stats": st.as_dict()}
        if record:
            res["replay_offsets"] = keep[0]
            res["replay_values"] = keep[1][: opts.record_count].copy()
        if debug:
            res["debug_hits"] = dbg.reshape(H * W * params.spp, params.max_depth, 3)
        return res


def render_params(spp=16, max_depth=5, rr_threshold=1.0, seed=0, **kw):
    p = abi.rbw_render_params(spp=spp, max_depth=max_depth, rr_threshold=rr_threshold, seed=seed)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def camera_rays(cam, raster_xy, lens_uv=None):
    """perspective_camera::sample for raster positions (n x 2) and lens samples (n x 2):
    world (origins, directions), n x 3 each"""
    xy = np.ascontiguousarray(raster_xy, np.float32).reshape(-1, 2)
    uv = np.zeros_like(xy) if lens_uv is None else np.ascontiguousarray(lens_uv, np.float32).reshape(-1, 2)
    o = np.zeros((len(xy), 3), np.float32)
    d = np.zeros((len(xy), 3), np.float32)
    lib().orc_camera_rays(C.byref(cam), len(xy), abi.ptr(xy, abi.f32p), abi.ptr(uv, abi.f32p), abi.ptr(o, abi.f32p),
                          abi.ptr(d, abi.f32p))
    return o, d


def pcg32_reals(seed, n):
    out = np.zeros(n, np.float32)
    lib().orc_pcg32_reals(seed, n, abi.ptr(out, abi.f32p))
    return out


def pcg32_range(seed, lo, hi):
    lo = np.ascontiguousarray(lo, np.uint32)
    hi = np.ascontiguousarray(hi, np.uint32)
    out = np.zeros(len(lo), np.uint32)
    lib().orc_pcg32_range(seed, len(lo), abi.ptr(lo, abi.u32p), abi.ptr(hi, abi.u32p), abi.ptr(out, abi.u32p))
    return out


def stratified_tables(seed, strata, dims, skip=0):
    """one pixel's reset(): (1D tables dims_1d x n, 2D tables dims_2d x n x 2, draws made)"""
    n = strata[0] * strata[1]
    t1 = np.zeros((dims[0], n), np.float32)
    t2 = np.zeros((dims[1], n, 2), np.float32)
    draws = lib().orc_stratified_tables(seed, skip, strata[0], strata[1], dims[0], dims[1], abi.ptr(t1, abi.f32p),
                                        abi.ptr(t2, abi.f32p))
    return t1, t2, int(draws)


def pcg32_u32(seed, n):
    out = np.zeros(n, np.uint32)
    lib().orc_pcg32_u32(seed, n, abi.ptr(out, abi.u32p))
    return out
