"""Scene and camera builders for the path-integrator workloads.

A scene is flattened exactly as rainbow's entity list would be
(scenes/scene.cpp:15-28): one entity per mesh or sphere, materials as
constant textures, area lights as surface_light on a mesh with a black matte
material (an emitter without a material is invisible to shadow rays,
scenes/entity.cpp:80-83, and is passed through by integrator.cpp:355-363).

Workloads (sizes as stated in the build's benchmark list):
  cornell_box()        Cornell-style box, 5 Lambert walls as 10 triangles, 2 spheres
                       (Lambert 0.8 / GGX conductor alpha 0.1), 0.25x0.25 quad light
                       radiance (17,12,4); pinhole 40 deg, 256x256
  displaced_icosphere_box(level=8)
                       the box plus a seeded fBm-displaced icosphere (20*4^level
                       triangles), per-triangle material: 50% Lambert albedo U[0.2,0.9],
                       else GGX conductor alpha U[0.05,0.5]
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

from . import abi

# numpy layout of rbw_material (13 packed 4-byte fields)
MATERIAL_DTYPE = np.dtype([("type", np.uint32), ("diffuse", np.float32, 3), ("sigma", np.float32),
                           ("eta", np.float32, 3), ("k", np.float32, 3), ("alpha_u", np.float32),
                           ("alpha_v", np.float32), ("specular", np.float32, 3), ("ior", np.float32)])
assert MATERIAL_DTYPE.itemsize == C.sizeof(abi.rbw_material)

# gold conductor (pbrt-v3 Au, RGB)
GOLD_ETA = (0.143119, 0.374957, 1.442479)
GOLD_K = (3.983160, 2.385721, 1.603215)


@dataclass
class Scene:
    positions: list = field(default_factory=list)   # list of (n,3) float32 blocks
    normals: list = field(default_factory=list)
    indices: list = field(default_factory=list)     # (m,3) uint32 global indices
    tri_entity: list = field(default_factory=list)
    tri_material: list = field(default_factory=list)
    spheres: list = field(default_factory=list)     # (cx, cy, cz, r)
    sphere_entity: list = field(default_factory=list)
    entities: list = field(default_factory=list)    # dicts
    materials: list = field(default_factory=list)   # dicts
    media: list = field(default_factory=list)      # abi.rbw_medium
    grid: list = field(default_factory=list)       # float32 blocks of grid densities
    nv: int = 0
    nt: int = 0
    has_normals: bool = False
    _keep: list = field(default_factory=list)

    # -- media (media/medium.hpp) --------------------------------------------
    def homogeneous_medium(self, sigma_a, sigma_s, g=0.0):
        """homogeneous_medium (media/homogeneous_medium.cpp:7-10)"""
        m = abi.rbw_medium(type=abi.RBW_MEDIUM_HOMOGENEOUS, g=g)
        m.sigma_a[:] = [float(v) for v in np.broadcast_to(sigma_a, 3)]
        m.sigma_s[:] = [float(v) for v in np.broadcast_to(sigma_s, 3)]
        self.media.append(m)
        return len(self.media) - 1

    def grid_medium(self, density, albedo, g=0.0, box_min=(0, 0, 0), box_max=(1, 1, 1)):
        """heterogeneous_medium (media/heterogeneous_medium.cpp:5-22): density[z, y, x] (sigma_t)
        over the axis-aligned box [box_min, box_max], trilinear; constant albedo and g"""
        density = np.ascontiguousarray(density, np.float32)
        nz, ny, nx = density.shape
        off = int(sum(len(b) for b in self.grid))
        self.grid.append(density.reshape(-1))
        lo, hi = np.asarray(box_min, np.float64), np.asarray(box_max, np.float64)
        sc = 1.0 / (hi - lo)
        w2l = [sc[0], 0, 0, -lo[0] * sc[0], 0, sc[1], 0, -lo[1] * sc[1], 0, 0, sc[2], -lo[2] * sc[2]]
        m = abi.rbw_medium(type=abi.RBW_MEDIUM_GRID, g=g, grid_offset=off)
        m.albedo[:] = [float(v) for v in np.broadcast_to(albedo, 3)]
        m.grid_size[:] = [nx, ny, nz]
        m.world_to_local[:] = [float(v) for v in w2l]
        self.media.append(m)
        return len(self.media) - 1

    def set_media(self, entity, outside=None, inside=None):
        """the entity's media component (media/medium.hpp:60-75): None = vacuum"""
        e = self.entities[entity]
        e["flags"] |= abi.RBW_ENTITY_HAS_MEDIA
        e["media"] = (abi.RBW_NONE if outside is None else outside, abi.RBW_NONE if inside is None else inside)

    # -- materials ---------------------------------------------------------
    def matte(self, rgb):
        m = np.zeros(1, MATERIAL_DTYPE)
        m["type"] = abi.RBW_MATERIAL_MATTE
        m["diffuse"] = rgb
        self.materials.append(m)
        return self.num_materials() - 1

    def metal(self, alpha, eta=GOLD_ETA, k=GOLD_K):
        m = np.zeros(1, MATERIAL_DTYPE)
        m["type"] = abi.RBW_MATERIAL_METAL
        m["eta"], m["k"], m["alpha_u"], m["alpha_v"] = eta, k, alpha, alpha
        self.materials.append(m)
        return self.num_materials() - 1

    def plastic(self, diffuse, specular, alpha, ior=1.49):
        """plastic_material (plastic_material.cpp:30-56): Lambert + GGX with dielectric Fresnel"""
        m = np.zeros(1, MATERIAL_DTYPE)
        m["type"] = abi.RBW_MATERIAL_PLASTIC
        m["diffuse"], m["specular"], m["alpha_u"], m["alpha_v"], m["ior"] = diffuse, specular, alpha, alpha, ior
        self.materials.append(m)
        return self.num_materials() - 1

    def add_materials(self, block):
        first = self.num_materials()
        self.materials.append(np.ascontiguousarray(block, MATERIAL_DTYPE))
        return first

    def num_materials(self):
        return int(sum(len(m) for m in self.materials))

    # -- entities ----------------------------------------------------------
    def add_mesh(self, positions, indices, material, normals=None, radiance=None, tri_material=None):
        positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        indices = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1, 3)
        eid = len(self.entities)
        flags = 0
        if normals is not None:
            flags |= abi.RBW_ENTITY_HAS_NORMALS
            self.has_normals = True
            normals = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        else:
            normals = np.zeros_like(positions)
        if radiance is not None:
            flags |= abi.RBW_ENTITY_EMITTER
        self.entities.append(dict(shape=abi.RBW_SHAPE_MESH, first=self.nt, count=len(indices),
                                  material=abi.RBW_NONE if material is None else material, flags=flags,
                                  radiance=tuple(radiance) if radiance is not None else (0.0, 0.0, 0.0)))
        self.positions.append(positions)
        self.normals.append(normals)
        self.indices.append(indices + np.uint32(self.nv))
        self.tri_entity.append(np.full(len(indices), eid, np.uint32))
        if tri_material is None:
            tri_material = np.full(len(indices), abi.RBW_NONE, np.uint32)
        self.tri_material.append(np.ascontiguousarray(tri_material, dtype=np.uint32))
        self.nv += len(positions)
        self.nt += len(indices)
        return eid

    def add_sphere(self, center, radius, material):
        eid = len(self.entities)
        sid = len(self.spheres)
        self.entities.append(dict(shape=abi.RBW_SHAPE_SPHERE, first=sid, count=1, material=material, flags=0,
                                  radiance=(0.0, 0.0, 0.0)))
        self.spheres.append((float(center[0]), float(center[1]), float(center[2]), float(radius)))
        self.sphere_entity.append(eid)
        return eid

    # -- flattening --------------------------------------------------------
    def arrays(self):
        P = np.concatenate(self.positions).astype(np.float32) if self.positions else np.zeros((0, 3), np.float32)
        N = np.concatenate(self.normals).astype(np.float32) if self.normals else np.zeros((0, 3), np.float32)
        I = np.concatenate(self.indices).astype(np.uint32) if self.indices else np.zeros((0, 3), np.uint32)
        TE = np.concatenate(self.tri_entity).astype(np.uint32) if self.tri_entity else np.zeros(0, np.uint32)
        TM = np.concatenate(self.tri_material).astype(np.uint32) if self.tri_material else np.zeros(0, np.uint32)
        S = np.array(self.spheres, np.float32).reshape(-1, 4)
        SE = np.array(self.sphere_entity, np.uint32)
        return P, N, I, TE, TM, S, SE

    def desc(self) -> abi.rbw_scene_desc:
        P, N, I, TE, TM, S, SE = self.arrays()
        P, N, I, TE, TM, S, SE = (np.ascontiguousarray(a) for a in (P, N, I, TE, TM, S, SE))
        ents = (abi.rbw_entity * len(self.entities))()
        for i, e in enumerate(self.entities):
            ents[i].shape, ents[i].first, ents[i].count = e["shape"], e["first"], e["count"]
            ents[i].material, ents[i].flags = e["material"], e["flags"]
            ents[i].radiance[:] = e["radiance"]
        mats = np.ascontiguousarray(np.concatenate(self.materials)) if self.materials else np.zeros(1, MATERIAL_DTYPE)
        use_tm = bool(len(TM)) and bool(np.any(TM != abi.RBW_NONE))
        d = abi.rbw_scene_desc(
            num_vertices=len(P), positions=abi.ptr(P, abi.f32p),
            normals=abi.ptr(N, abi.f32p) if self.has_normals else abi.ptr(None, abi.f32p),
            uvs=abi.ptr(None, abi.f32p),
            num_triangles=len(I), indices=abi.ptr(I, abi.u32p), triangle_entity=abi.ptr(TE, abi.u32p),
            triangle_material=abi.ptr(TM, abi.u32p) if use_tm else abi.ptr(None, abi.u32p),
            num_spheres=len(S), spheres=abi.ptr(S, abi.f32p) if len(S) else abi.ptr(None, abi.f32p),
            sphere_entity=abi.ptr(SE, abi.u32p) if len(S) else abi.ptr(None, abi.u32p),
            num_entities=len(self.entities), entities=ents, num_materials=self.num_materials(),
            materials=mats.ctypes.data_as(C.POINTER(abi.rbw_material)))
        self._keep = [P, N, I, TE, TM, S, SE, ents, mats]
        if self.media:
            med = (abi.rbw_medium * len(self.media))(*self.media)
            em = np.array([e.get("media", (abi.RBW_NONE, abi.RBW_NONE)) for e in self.entities], np.uint32).reshape(-1)
            G = np.concatenate(self.grid).astype(np.float32) if self.grid else np.zeros(1, np.float32)
            d.num_media, d.media = len(self.media), med
            d.entity_media = abi.ptr(em, abi.u32p)
            d.medium_grid_count = int(sum(len(b) for b in self.grid))
            d.medium_grid = abi.ptr(G, abi.f32p)
            self._keep += [med, em, G]
        return d


# ------------------------------------------------------------------ camera

def _perspective_fov_rh(fov, width, height, near=0.01, far=1000.0):
    """glm::perspectiveFovRH_NO (column-major, float64)."""
    h = math.cos(0.5 * fov) / math.sin(0.5 * fov)
    w = h * height / width
    m = np.zeros((4, 4))
    m[0, 0] = w
    m[1, 1] = h
    m[2, 2] = -(far + near) / (far - near)
    m[2, 3] = -1.0
    m[3, 2] = -(2.0 * far * near) / (far - near)
    return m  # m[col, row]


def _scale(v):
    m = np.eye(4)
    m[0, 0], m[1, 1], m[2, 2] = v
    return m


def _translate(v):
    m = np.eye(4)
    m[3, 0:3] = v
    return m


def _look_at_rh(eye, center, up):
    """glm::lookAtRH (column-major, float64)."""
    eye, center, up = (np.asarray(x, float) for x in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, 0], m[1, 0], m[2, 0] = s
    m[0, 1], m[1, 1], m[2, 1] = u
    m[0, 2], m[1, 2], m[2, 2] = -f
    m[3, 0] = -np.dot(s, eye)
    m[3, 1] = -np.dot(u, eye)
    m[3, 2] = np.dot(f, eye)
    return m


def _mul(a, b):
    """column-major product a*b (stored m[col,row])."""
    return (a.T @ b.T).T


def _inv(a):
    return np.linalg.inv(a.T).T


def perspective_camera(width, height, fov_deg, eye, target, up=(0, 1, 0), lens_radius=0.0, focus=1e6):
    """perspective_camera (cameras/camera.cpp:21-52, perspective_camera.cpp:1-20) with the
    screen window [-1,1]^2; matrices composed in float64 and rounded.  The projection
    (glm's perspectiveFovRH, shared/transform.cpp:144) already scales x by height/width,
    so this window gives square pixels: fov_deg is the vertical field of view and the
    horizontal half-angle is atan(tan(fov/2) * width/height)."""
    sw_min, sw_max = (-1.0, -1.0), (1.0, 1.0)
    # the reference's perspective_right_hand takes float fov/near/far (shared/transform.hpp:113)
    fov = float(np.float32(math.radians(fov_deg)))
    cam_to_screen = _perspective_fov_rh(fov, float(np.float32(width)), float(np.float32(height)),
                                        near=float(np.float32(0.01)), far=1000.0)
    s2r = _scale((width, height, 1.0))
    s2r = _mul(s2r, _scale((1.0 / (sw_max[0] - sw_min[0]), 1.0 / (sw_min[1] - sw_max[1]), 1.0)))
    s2r = _mul(s2r, _translate((-sw_min[0], -sw_max[1], 0.0)))
    r2c = _mul(_inv(cam_to_screen), _inv(s2r))
    r2c[np.abs(r2c) < 1e-12] = 0.0  # inversion residue of entries that are exactly zero
    c2w = _inv(_look_at_rh(eye, target, up))
    cam = abi.rbw_camera_desc()
    cam.raster_to_camera[:] = [float(x) for x in r2c.astype(np.float32).reshape(-1)]
    cam.camera_to_world[:] = [float(x) for x in c2w.astype(np.float32).reshape(-1)]
    cam.width, cam.height = width, height
    cam.bound_min[:] = (0, 0)
    cam.bound_max[:] = (width, height)
    cam.lens_radius = lens_radius
    cam.focus_distance = focus
    return cam


# ------------------------------------------------------------------ workloads

def _quad(p0, p1, p2, p3):
    """two triangles (p0,p1,p2), (p0,p2,p3) with the face normal on every vertex."""
    P = np.array([p0, p1, p2, p3], np.float32)
    n = np.cross(P[1] - P[0], P[2] - P[0])
    n = n / np.linalg.norm(n)
    N = np.tile(n.astype(np.float32), (4, 1))
    I = np.array([[0, 1, 2], [0, 2, 3]], np.uint32)
    return P, I, N


def add_cornell_box(sc: Scene, light_size=0.25, radiance=(17.0, 12.0, 4.0), spheres=True):
    white = sc.matte((0.73, 0.73, 0.73))
    red = sc.matte((0.65, 0.05, 0.05))
    green = sc.matte((0.12, 0.45, 0.15))
    black = sc.matte((0.0, 0.0, 0.0))
    # walls with inward normals (counter-clockwise seen from inside)
    walls = [
        (_quad((-1, -1, 1), (1, -1, 1), (1, -1, -1), (-1, -1, -1)), white),   # floor y=-1, n=+y
        (_quad((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1)), white),       # ceiling y=+1, n=-y
        (_quad((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1)), white),   # back z=-1, n=+z
        (_quad((-1, -1, 1), (-1, -1, -1), (-1, 1, -1), (-1, 1, 1)), red),     # left x=-1, n=+x
        (_quad((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1)), green),       # right x=+1, n=-x
    ]
    for (P, I, N), m in walls:
        sc.add_mesh(P, I, m, normals=N)
    h = light_size / 2
    y = 0.999
    P, I, N = _quad((-h, y, -h), (h, y, -h), (h, y, h), (-h, y, h))  # facing down
    light = sc.add_mesh(P, I, black, normals=N, radiance=radiance)
    if spheres:
        sc.add_sphere((-0.45, -0.7, -0.35), 0.3, sc.matte((0.8, 0.8, 0.8)))
        sc.add_sphere((0.5, -0.8, 0.3), 0.2, sc.metal(0.1))
    return light


def cornell_box(width=256, height=256):
    """Cornell-style box, 10 wall triangles, 2 spheres, quad light; returns (scene, camera)."""
    sc = Scene()
    add_cornell_box(sc)
    cam = perspective_camera(width, height, 40.0, eye=(0.0, 0.0, 3.5), target=(0.0, 0.0, 0.0))
    return sc, cam


def box_mesh(lo, hi):
    """an axis-aligned box of 12 triangles with outward face normals: (P, I, N)"""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = [((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)),   # +z
             ((x1, y0, z0), (x0, y0, z0), (x0, y1, z0), (x1, y1, z0)),   # -z
             ((x1, y0, z1), (x1, y0, z0), (x1, y1, z0), (x1, y1, z1)),   # +x
             ((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0)),   # -x
             ((x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (x0, y1, z0)),   # +y
             ((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1))]   # -y
    Ps, Is, Ns = [], [], []
    for k, f in enumerate(faces):
        P, I, N = _quad(*f)
        Ps.append(P)
        Is.append(I + 4 * k)
        Ns.append(N)
    return np.concatenate(Ps), np.concatenate(Is), np.concatenate(Ns)


def smoke_grid(n=16, seed=3):
    """a soft blob of density on an n^3 grid (density[z, y, x], sigma_t up to ~6)"""
    rng = np.random.default_rng(seed)
    c = (np.arange(n) + 0.5) / n
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    r2 = (x - 0.5) ** 2 + (y - 0.45) ** 2 + (z - 0.5) ** 2
    d = 6.0 * np.exp(-r2 / 0.06) * (0.75 + 0.5 * rng.random((n, n, n)))
    return d.astype(np.float32)


def foggy_cornell_box(width=256, height=256, grid=True):
    """The Cornell box with participating media for the volume path integrator: an
    invisible sphere holding a homogeneous forward-scattering medium and, when `grid`,
    an invisible box holding a density grid; camera outside both (it starts in vacuum)."""
    sc = Scene()
    add_cornell_box(sc, spheres=False)
    fog = sc.homogeneous_medium(sigma_a=0.2, sigma_s=(1.6, 1.2, 0.9), g=0.5)
    e = sc.add_sphere((-0.4, -0.5, -0.25), 0.45, abi.RBW_NONE)
    sc.set_media(e, inside=fog)
    if grid:
        # the box stops short of the floor: coplanar faces would make the closest hit a tie
        smoke = sc.grid_medium(smoke_grid(), albedo=(0.9, 0.85, 0.8), g=-0.2, box_min=(0.05, -0.99, -0.6),
                               box_max=(0.9, 0.1, 0.4))
        P, I, N = box_mesh((0.05, -0.99, -0.6), (0.9, 0.1, 0.4))
        e = sc.add_mesh(P, I, None, normals=N)
        sc.set_media(e, inside=smoke)
    cam = perspective_camera(width, height, 40.0, eye=(0.0, 0.0, 3.5), target=(0.0, 0.0, 0.0))
    return sc, cam


def _pcg32_stream(seed, n):
    """pcg32 (random_generator.cpp:41-66) uniform floats in numpy, sequential."""
    mult = 0x5851f42d4c957f2d
    mask = (1 << 64) - 1
    state = 0
    inc = ((seed << 1) | 1) & mask

    def step(st):
        old = st
        st = (old * mult + inc) & mask
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return st, ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    state, _ = step(state)
    state = (state + 0x853c49e6748fea9b) & mask
    state, _ = step(state)
    out = np.empty(n, np.float32)
    for i in range(n):
        state, v = step(state)
        out[i] = min(np.float32(v) * np.float32(2.3283064365386963e-10), np.float32(1 - 2 ** -23))
    return out


def icosphere(level):
    """unit icosphere subdivided `level` times: (vertices float64 (V,3), faces uint32 (F,3))."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], float)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                  (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                  (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], np.int64)
    for _ in range(level):
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (len(V) + 1) + e[:, 1]
        uk, inv = np.unique(key, return_inverse=True)
        a = uk // (len(V) + 1)
        b = uk % (len(V) + 1)
        mid = V[a] + V[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = len(V)
        V = np.concatenate([V, mid])
        nf = len(F)
        m01 = base + inv[:nf]
        m12 = base + inv[nf:2 * nf]
        m20 = base + inv[2 * nf:]
        F = np.concatenate([np.stack([F[:, 0], m01, m20], 1), np.stack([F[:, 1], m12, m01], 1),
                            np.stack([F[:, 2], m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return V, F.astype(np.uint32)


def _value_noise3(p, table):
    """trilinear value noise on a hashed lattice; p (n,3) float64, table (256,) of floats in [0,1)."""
    perm = (np.argsort(table[:256]) & 255).astype(np.int64)
    vals = table.astype(np.float64)
    pi = np.floor(p).astype(np.int64)
    f = p - pi
    f = f * f * (3 - 2 * f)

    def h(ix, iy, iz):
        return vals[perm[(perm[(perm[ix & 255] + iy) & 255] + iz) & 255]]

    x, y, z = pi[:, 0], pi[:, 1], pi[:, 2]
    c = {}
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                c[(dx, dy, dz)] = h(x + dx, y + dy, z + dz)
    fx, fy, fz = f[:, 0], f[:, 1], f[:, 2]
    x00 = c[(0, 0, 0)] * (1 - fx) + c[(1, 0, 0)] * fx
    x10 = c[(0, 1, 0)] * (1 - fx) + c[(1, 1, 0)] * fx
    x01 = c[(0, 0, 1)] * (1 - fx) + c[(1, 0, 1)] * fx
    x11 = c[(0, 1, 1)] * (1 - fx) + c[(1, 1, 1)] * fx
    y0 = x00 * (1 - fy) + x10 * fy
    y1 = x01 * (1 - fy) + x11 * fy
    return y0 * (1 - fz) + y1 * fz


def fbm_displaced_icosphere(level=8, seed=1, center=(0.05, -0.35, -0.3), radius=0.42, amplitude=0.1, octaves=6):
    """icosphere with vertices displaced along the normal by seeded fBm; smooth vertex normals."""
    V, F = icosphere(level)
    table = _pcg32_stream(seed, 256)
    n = np.zeros(len(V))
    freq, amp = 3.0, 1.0
    for _ in range(octaves):
        n += amp * (_value_noise3(V * freq + 17.0, table) - 0.5)
        freq *= 2.0
        amp *= 0.5
    P = V * (radius * (1.0 + amplitude * n))[:, None] + np.asarray(center)
    P32 = P.astype(np.float32)
    Pd = P32.astype(np.float64)
    fn = np.cross(Pd[F[:, 1]] - Pd[F[:, 0]], Pd[F[:, 2]] - Pd[F[:, 0]])
    N = np.zeros_like(Pd)
    for k in range(3):
        np.add.at(N, F[:, k], fn)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    return P32, F, N.astype(np.float32)


def per_triangle_materials(sc: Scene, ntri, seed, alpha_range=(0.05, 0.5)):
    """50% Lambert albedo U[0.2,0.9]^3, else GGX conductor alpha U[alpha_range]; one material per triangle."""
    rng = np.random.default_rng(seed)
    coin = rng.random(ntri) < 0.5
    albedo = rng.uniform(0.2, 0.9, (ntri, 3))
    alpha = rng.uniform(alpha_range[0], alpha_range[1], ntri)
    m = np.zeros(ntri, MATERIAL_DTYPE)
    m["type"] = np.where(coin, abi.RBW_MATERIAL_MATTE, abi.RBW_MATERIAL_METAL)
    m["diffuse"] = np.where(coin[:, None], albedo, 0.0)
    m["eta"] = np.where(coin[:, None], 1.0, np.asarray(GOLD_ETA))
    m["k"] = np.where(coin[:, None], 0.0, np.asarray(GOLD_K))
    m["alpha_u"] = np.where(coin, 0.0, alpha)
    m["alpha_v"] = m["alpha_u"]
    first = sc.add_materials(m)
    return np.arange(first, first + ntri, dtype=np.uint32)


def displaced_icosphere_box(level=8, width=1024, height=1024, seed=1):
    """the box plus an fBm-displaced icosphere (20*4^level triangles) with per-triangle materials."""
    sc = Scene()
    add_cornell_box(sc)
    P, F, N = fbm_displaced_icosphere(level=level, seed=seed)
    tm = per_triangle_materials(sc, len(F), seed)
    sc.add_mesh(P, F, tm[0], normals=N, tri_material=tm)
    cam = perspective_camera(width, height, 40.0, eye=(0.0, 0.0, 3.5), target=(0.0, 0.0, 0.0))
    return sc, cam


def fbm_displaced_grid_sphere(ntri, seed=1, center=(0.05, -0.35, -0.3), radius=0.42, amplitude=0.1, octaves=6):
    """latitude-longitude grid sphere of about `ntri` triangles (pole fans plus quad rings),
    vertices displaced along the normal by seeded fBm; smooth vertex normals."""
    nr = max(2, int(round((ntri / 4.0) ** 0.5)) + 1)       # rings of latitude (including poles)
    ns = 2 * (nr - 1)                                       # segments of longitude
    th = np.linspace(0.0, math.pi, nr + 1)[1:-1]            # interior rings
    ph = np.linspace(0.0, 2 * math.pi, ns, endpoint=False)
    ring = np.stack([np.outer(np.sin(th), np.cos(ph)), np.repeat(np.cos(th)[:, None], ns, 1),
                     np.outer(np.sin(th), np.sin(ph))], -1).reshape(-1, 3)
    V = np.concatenate([[(0.0, 1.0, 0.0)], ring, [(0.0, -1.0, 0.0)]])
    south = len(V) - 1
    F = []
    j = np.arange(ns)
    jn = (j + 1) % ns
    F.append(np.stack([np.zeros(ns, np.int64), 1 + jn, 1 + j], 1))
    for r in range(nr - 2):
        a, b = 1 + r * ns, 1 + (r + 1) * ns
        F.append(np.stack([a + j, a + jn, b + jn], 1))
        F.append(np.stack([a + j, b + jn, b + j], 1))
    last = 1 + (nr - 2) * ns
    F.append(np.stack([last + j, last + jn, np.full(ns, south)], 1))
    F = np.concatenate(F).astype(np.int64)
    table = _pcg32_stream(seed, 256)
    n = np.zeros(len(V))
    freq, amp = 3.0, 1.0
    for _ in range(octaves):
        n += amp * (_value_noise3(V * freq + 17.0, table) - 0.5)
        freq *= 2.0
        amp *= 0.5
    P = V * (radius * (1.0 + amplitude * n))[:, None] + np.asarray(center)
    P32 = P.astype(np.float32)
    Pd = P32.astype(np.float64)
    fn = np.cross(Pd[F[:, 1]] - Pd[F[:, 0]], Pd[F[:, 2]] - Pd[F[:, 0]])
    N = np.zeros_like(Pd)
    for k in range(3):
        np.add.at(N, F[:, k], fn)
    N /= np.maximum(np.linalg.norm(N, axis=1, keepdims=True), 1e-30)
    return P32, F.astype(np.uint32), N.astype(np.float32)


def batch_workload(n_jobs=64, seed=2, tris=(1e4, 2e6), res=(128, 1024), spp=(4, 256), depth=(3, 12)):
    """Jobs of uneven size: per job a seeded fBm-displaced grid sphere with per-triangle
    Lambert/GGX materials inside the Cornell box; triangle count, square resolution and
    spp log-uniform in their ranges, max depth uniform in [depth[0], depth[1]].
    Returns a list of dicts {scene, camera, spp, max_depth, seed, triangles}."""
    rng = np.random.default_rng(seed)
    params = []
    for i in range(n_jobs):
        t = int(round(math.exp(rng.uniform(math.log(tris[0]), math.log(tris[1])))))
        w = int(round(math.exp(rng.uniform(math.log(res[0]), math.log(res[1])))))
        s = int(round(math.exp(rng.uniform(math.log(spp[0]), math.log(spp[1])))))
        d = int(rng.integers(depth[0], depth[1] + 1))
        params.append((i, t, w, s, d))

    def build(prm):
        i, t, w, s, d = prm
        sc = Scene()
        add_cornell_box(sc, spheres=False)
        P, F, N = fbm_displaced_grid_sphere(t, seed=1000 + i)
        tm = per_triangle_materials(sc, len(F), 1000 + i)
        sc.add_mesh(P, F, tm[0], normals=N, tri_material=tm)
        cam = perspective_camera(w, w, 40.0, eye=(0.0, 0.0, 3.5), target=(0.0, 0.0, 0.0))
        return {"scene": sc, "camera": cam, "spp": s, "max_depth": d, "seed": 7 + i, "triangles": len(F) + 10}

    # each job's scene depends only on its own parameters and seed: build them on threads
    # (numpy's array loops run outside the interpreter lock), largest first
    from concurrent.futures import ThreadPoolExecutor
    order = sorted(range(n_jobs), key=lambda k: -params[k][1])
    with ThreadPoolExecutor(max(1, min(16, os.cpu_count() or 1))) as ex:
        built = dict(zip(order, ex.map(build, [params[k] for k in order])))
    return [built[k] for k in range(n_jobs)]


def narrow_aperture_room(level=8, width=512, height=512, gap=0.3, radiance=50.0, seed=1, mesh=True,
                         light_dist=0.5, light_size=0.5, light_tilt=0.0):
    """A closed room lit only through a gap x gap hole in its +x wall by a quad light outside
    (radiance 50, light_size square, light_dist from the hole centre along its normal, which
    points through the hole tilted light_tilt degrees upward from -x); the displaced
    icosphere with per-triangle materials inside; camera inside."""
    sc = Scene()
    white = sc.matte((0.73, 0.73, 0.73))
    black = sc.matte((0.0, 0.0, 0.0))
    walls = [((-1, -1, 1), (1, -1, 1), (1, -1, -1), (-1, -1, -1)),     # floor, n=+y
             ((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1)),         # ceiling, n=-y
             ((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1)),     # back z=-1, n=+z
             ((1, -1, 1), (-1, -1, 1), (-1, 1, 1), (1, 1, 1)),         # front z=+1, n=-z
             ((-1, -1, 1), (-1, -1, -1), (-1, 1, -1), (-1, 1, 1))]     # left x=-1, n=+x
    for q in walls:
        P, I, N = _quad(*q)
        sc.add_mesh(P, I, white, normals=N)
    # +x wall with a hole centred at (y, z) = (cy, 0)
    cy, h = 0.2, gap / 2
    y0, y1, z0, z1 = cy - h, cy + h, -h, h
    rects = [(-1, y0, -1, 1), (y1, 1, -1, 1), (y0, y1, -1, z0), (y0, y1, z1, 1)]  # (ylo, yhi, zlo, zhi)
    Ps, Is, Ns = [], [], []
    for k, (ya, yb, za, zb) in enumerate(rects):
        P, I, N = _quad((1, ya, za), (1, ya, zb), (1, yb, zb), (1, yb, za))   # n=-x
        Ps.append(P)
        Is.append(I + 4 * k)
        Ns.append(N)
    sc.add_mesh(np.concatenate(Ps), np.concatenate(Is), white, normals=np.concatenate(Ns))
    # the light outside, its normal through the hole centre
    th = math.radians(light_tilt)
    n = np.array([-math.cos(th), math.sin(th), 0.0])
    u = np.array([math.sin(th), math.cos(th), 0.0])
    v = np.array([0.0, 0.0, 1.0])
    c = np.array([1.0, cy, 0.0]) - light_dist * n
    s = 0.5 * light_size
    corners = [c - s * u - s * v, c - s * u + s * v, c + s * u + s * v, c + s * u - s * v]
    if np.dot(np.cross(corners[1] - corners[0], corners[2] - corners[0]), n) < 0:
        corners = corners[::-1]
    P, I, N = _quad(*corners)
    sc.add_mesh(P, I, black, normals=N, radiance=(radiance,) * 3)
    if mesh:
        P, F, Nn = fbm_displaced_icosphere(level=level, seed=seed, center=(-0.2, -0.45, -0.3), radius=0.45)
        tm = per_triangle_materials(sc, len(F), seed)
        sc.add_mesh(P, F, tm[0], normals=Nn, tri_material=tm)
    cam = perspective_camera(width, height, 65.0, eye=(0.6, 0.3, 0.95), target=(-0.3, -0.3, -0.6))
    return sc, cam


def room_variants(n=16, seed=4, res=256, tris=(1e5, 1.3e6)):
    """Variants of the narrow-aperture room: gap U[0.1, 0.5], a seeded fBm-displaced grid
    sphere of tris[0]-tris[1] triangles (log-uniform) with per-triangle Lambert / GGX
    materials (alpha U[0.02, 0.6]), the light aimed 45 degrees up through the gap; res x res,
    or, for res = (lo, hi), a square resolution log-uniform in [lo, hi] per variant (drawn
    from a generator of its own, so gaps and meshes do not depend on it).
    Returns a list of dicts {scene, camera, gap, triangles, seed, res}."""
    rng = np.random.default_rng(seed)
    rres = np.random.default_rng(seed + 1000)
    out = []
    for i in range(n):
        gap = float(rng.uniform(0.1, 0.5))
        t = int(round(math.exp(rng.uniform(math.log(tris[0]), math.log(tris[1])))))
        r = res if np.isscalar(res) else int(round(math.exp(rres.uniform(math.log(res[0]), math.log(res[1])))))
        sc, cam = narrow_aperture_room(level=8, width=r, height=r, gap=gap, light_tilt=45.0, mesh=False)
        P, F, N = fbm_displaced_grid_sphere(t, seed=2000 + i, center=(-0.2, -0.45, -0.3), radius=0.45)
        tm = per_triangle_materials(sc, len(F), 2000 + i, alpha_range=(0.02, 0.6))
        sc.add_mesh(P, F, tm[0], normals=N, tri_material=tm)
        out.append({"scene": sc, "camera": cam, "gap": gap, "triangles": len(F), "seed": 2000 + i, "res": r})
    return out


def rel_mse(img, ref):
    """mean over pixels and channels of (I - R)^2 / (R^2 + 0.01)"""
    return float(np.mean((img.astype(np.float64) - ref) ** 2 / (ref.astype(np.float64) ** 2 + 0.01)))
