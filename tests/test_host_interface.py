"""The rainbow-style C++ host interface (rainbow_amd/host): a scene built with
entities, meshes, spheres, materials and a perspective camera flattens to the
same C-ABI descriptors as the Python builders, and (on the GPU) renders the
same film through wavefront_path_integrator::render."""
import os
import subprocess

import numpy as np
import pytest

from rainbow_amd import abi, scenes
from rainbow_amd.build import TEST_BIN

MATERIAL_DTYPE = scenes.MATERIAL_DTYPE


def _read_flat(path):
    import ctypes as C
    b = open(path, "rb").read()
    off = 0

    def take(dtype, n):
        nonlocal off
        a = np.frombuffer(b, dtype, n, off)
        off += a.nbytes
        return a

    nv, nt, ns, ne, nm = take(np.uint32, 5)
    out = {"P": take(np.float32, 3 * nv), "N": take(np.float32, 3 * nv), "I": take(np.uint32, 3 * nt),
           "TE": take(np.uint32, nt), "S": take(np.float32, 4 * ns), "SE": take(np.uint32, ns)}
    out["E"] = b[off: off + ne * C.sizeof(abi.rbw_entity)]
    off += ne * C.sizeof(abi.rbw_entity)
    out["M"] = np.frombuffer(b, MATERIAL_DTYPE, nm, off)
    off += nm * MATERIAL_DTYPE.itemsize
    out["CAM"] = abi.rbw_camera_desc.from_buffer_copy(b[off: off + C.sizeof(abi.rbw_camera_desc)])
    off += C.sizeof(abi.rbw_camera_desc)
    has_uvs = int(take(np.uint32, 1)[0])
    out["UV"] = take(np.float32, 2 * nv) if has_uvs else None
    nmed = int(take(np.uint32, 1)[0])
    out["MED"] = None
    if nmed:
        out["MED"] = (abi.rbw_medium * nmed).from_buffer_copy(b[off: off + nmed * C.sizeof(abi.rbw_medium)])
        off += nmed * C.sizeof(abi.rbw_medium)
        out["EM"] = take(np.uint32, 2 * ne)
    ngrid = int(take(np.uint32, 1)[0])
    out["G"] = take(np.float32, ngrid) if ngrid else np.zeros(0, np.float32)
    assert off == len(b)
    return out


def flat_desc(f):
    """an rbw_scene_desc over the arrays of a flattened scene (the arrays stay in f)"""
    import ctypes as C
    d = abi.rbw_scene_desc()
    keep = {k: np.ascontiguousarray(v) for k, v in f.items() if isinstance(v, np.ndarray)}
    f["_keep"] = keep
    d.num_vertices = len(keep["P"]) // 3
    d.positions = abi.ptr(keep["P"], abi.f32p)
    d.normals = abi.ptr(keep["N"], abi.f32p)
    d.uvs = abi.ptr(keep["UV"], abi.f32p) if "UV" in keep else abi.ptr(None, abi.f32p)
    d.num_triangles = len(keep["I"]) // 3
    d.indices = abi.ptr(keep["I"], abi.u32p)
    d.triangle_entity = abi.ptr(keep["TE"], abi.u32p)
    d.num_spheres = len(keep["S"]) // 4
    d.spheres = abi.ptr(keep["S"], abi.f32p) if d.num_spheres else abi.ptr(None, abi.f32p)
    d.sphere_entity = abi.ptr(keep["SE"], abi.u32p) if d.num_spheres else abi.ptr(None, abi.u32p)
    ents = (abi.rbw_entity * (len(f["E"]) // C.sizeof(abi.rbw_entity))).from_buffer_copy(f["E"])
    mats = (abi.rbw_material * len(keep["M"])).from_buffer_copy(keep["M"].tobytes())
    f["_ents"], f["_mats"] = ents, mats
    d.num_entities = len(ents)
    d.entities = C.cast(ents, C.POINTER(abi.rbw_entity))
    d.num_materials = len(mats)
    d.materials = C.cast(mats, C.POINTER(abi.rbw_material))
    if f.get("MED") is not None:
        d.num_media = len(f["MED"])
        d.media = C.cast(f["MED"], C.POINTER(abi.rbw_medium))
        d.entity_media = abi.ptr(keep["EM"], abi.u32p)
        d.medium_grid_count = len(keep["G"])
        d.medium_grid = abi.ptr(keep["G"], abi.f32p) if len(keep["G"]) else abi.ptr(None, abi.f32p)
    return d


def _binary():
    if not os.path.exists(TEST_BIN):
        pytest.skip("C++ driver not built")
    return TEST_BIN


def test_host_flatten_matches_python_builder(tmp_path):
    import ctypes as C
    f = tmp_path / "flat.bin"
    subprocess.run([_binary(), "flatten", str(f)], check=True)
    got = _read_flat(f)
    sc, cam = scenes.cornell_box()
    P, N, I, TE, TM, S, SE = sc.arrays()
    assert np.array_equal(got["P"], P.reshape(-1))
    assert np.array_equal(got["N"], N.reshape(-1))
    assert np.array_equal(got["I"], I.reshape(-1))
    assert np.array_equal(got["TE"], TE)
    assert np.array_equal(got["S"], S.reshape(-1))
    assert np.array_equal(got["SE"], SE)
    d = sc.desc()
    ents = bytes((abi.rbw_entity * d.num_entities).from_address(C.addressof(d.entities.contents)))
    assert got["E"] == ents
    mats = np.concatenate(sc.materials)
    assert got["M"].tobytes() == mats.tobytes()
    gc = got["CAM"]
    assert np.allclose(np.array(gc.raster_to_camera), np.array(cam.raster_to_camera), rtol=2e-7, atol=0)
    assert list(gc.camera_to_world) == list(cam.camera_to_world)
    assert (gc.width, gc.height) == (cam.width, cam.height)


@pytest.mark.gpu
def test_host_integrator_render_matches_c_abi(tracer, tmp_path):
    """wavefront_path_integrator::render(camera, scene) == rbw_render on the Python-built scene."""
    from rainbow_amd.tracer import render_params
    f = tmp_path / "film.bin"
    subprocess.run([_binary(), "render", str(f), "16", "5", "3", str(tmp_path / "film")], check=True)
    img = np.fromfile(f, np.float32).reshape(256, 256, 3)
    # film::write: the PNG holds the film's 8-bit conversion
    from test_image import read_png
    from rainbow_amd.tracer import film_to_rgba8
    assert np.array_equal(read_png(tmp_path / "film.png"), film_to_rgba8(img))
    sc, cam = scenes.cornell_box()
    tracer.upload(sc)
    ref = tracer.render(cam, render_params(spp=16, max_depth=5, seed=3))["image"]
    bad = (np.abs(img - ref) > 1e-4 * np.maximum(np.abs(img), np.abs(ref)) + 1e-7).any(-1)
    assert bad.mean() < 1e-3, bad.sum()


@pytest.mark.gpu
def test_host_integrator_stratified_samplers(tracer, tmp_path):
    """stratified_sampler2d(4, 4, 12) and stratified_sampler1d(4, 4, 8) passed to the
    integrator select the GPU's stratified sampler with those strata and dimensions."""
    from rainbow_amd.tracer import render_params
    f = tmp_path / "film.bin"
    subprocess.run([_binary(), "render_stratified", str(f), "16", "5", "3"], check=True)
    img = np.fromfile(f, np.float32).reshape(256, 256, 3)
    sc, cam = scenes.cornell_box()
    tracer.upload(sc)
    ref = tracer.render(cam, render_params(spp=16, max_depth=5, seed=3, sampler=abi.RBW_SAMPLER_STRATIFIED,
                                           strata_x=4, strata_y=4, dims_1d=8, dims_2d=12))["image"]
    bad = (np.abs(img - ref) > 1e-4 * np.maximum(np.abs(img), np.abs(ref)) + 1e-7).any(-1)
    assert bad.mean() < 1e-3, bad.sum()
