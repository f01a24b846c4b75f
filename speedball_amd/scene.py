"""Scene assembly in the reference layouts through the native builder.

`SceneBuilder` wraps the sbgi_sb_* entry points (buildSpectralScene's BLAS/pool/TLAS
assembly, js/spectral_scene.js:898-1366). `Scene` holds the finalized arrays as numpy
copies plus a `view()` that points at them for sbgi_set_scene and for the oracle.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

MAT_STRIDE = 28
LIGHT_STRIDE = 17


def uber_material(color=(1.0, 1.0, 1.0), roughness=1.0, metalness=0.0, transmission=0.0, ior=1.5,
                  emissive=(0.0, 0.0, 0.0), opacity=1.0, side=0, alpha_test=0.0) -> np.ndarray:
    """materialToUber (js/spectral_scene.js:214-307) for a map-free material."""
    r = np.zeros(MAT_STRIDE, np.float32)
    r[0:3] = color
    r[3] = min(1.0, max(0.0, roughness))
    r[4] = min(1.0, max(0.0, metalness))
    tr = min(1.0, max(0.0, transmission))
    r[5] = tr
    r[6] = max(1.0, ior)
    r[7:10] = emissive
    r[10] = min(1.0, max(0.0, opacity))
    r[12:17] = -1
    r[17] = 1.0
    r[18:20] = 1.0
    r[22] = min(2, max(0, int(side)))
    r[23] = min(1.0, max(0.0, alpha_test))
    r[24] = -1
    r[25] = -1
    const_alpha_ok = opacity > 1.0e-4 and (alpha_test <= 0 or opacity >= alpha_test)
    flags = (2 if const_alpha_ok else 0) | (4 if tr < 0.5 else 0)
    r[26] = flags
    return r


def point_light(pos, color=(1.0, 1.0, 1.0), intensity=1.0, distance=0.0, decay=2.0) -> np.ndarray:
    """collectLights record for a PointLight (js/spectral_scene.js:754-813)."""
    r = np.zeros(LIGHT_STRIDE, np.float32)
    r[0] = 1
    r[1:4] = pos
    r[4:7] = (0, 0, -1)
    r[7:10] = np.asarray(color, np.float64) * intensity
    r[10] = distance
    r[11] = decay
    r[12] = -1
    r[13] = -1
    return r


def directional_light(pos, target, color=(1.0, 1.0, 1.0), intensity=1.0) -> np.ndarray:
    r = np.zeros(LIGHT_STRIDE, np.float32)
    d = np.asarray(target, np.float64) - np.asarray(pos, np.float64)
    d = d / np.linalg.norm(d)
    r[0] = 0
    r[1:4] = pos
    r[4:7] = d
    r[7:10] = np.asarray(color, np.float64) * intensity
    r[10] = 0
    r[11] = 2
    r[12] = -1
    r[13] = -1
    return r


def spot_light(pos, target, color=(1.0, 1.0, 1.0), intensity=1.0, distance=0.0, angle=math.pi / 4, penumbra=0.0,
               decay=2.0) -> np.ndarray:
    """collectLights record for a SpotLight (js/spectral_scene.js:785-797): the axis from
    the light to its target, cosAngle = cos(angle), cosPenumbra = cos(angle (1 - penumbra))."""
    r = directional_light(pos, target, color, intensity)
    r[0] = 2
    r[10] = distance
    r[11] = decay
    r[12] = math.cos(angle)
    r[13] = math.cos(angle * (1.0 - penumbra))
    return r


def emitter_light(center, radius: float, area: float, radiance) -> np.ndarray:
    """collectEmitterRecords record of an emissive mesh (js/spectral_scene.js:815-875):
    bounding-sphere centre and radius, surface radiance (emissive x intensity x scale) in
    the colour slots, the Cauchy projected area A/4 in [15] and the power-based range
    sqrt(max(radiance) A/4 / 0.005)."""
    rgb = np.asarray(radiance, np.float64)
    proj = float(area) / 4.0
    r = np.zeros(LIGHT_STRIDE, np.float32)
    r[0] = 3
    r[1:4] = center
    r[4:7] = (0, 0, -1)
    r[7:10] = rgb
    r[10] = np.sqrt(max(0.0, float(rgb.max())) * proj / 0.005)
    r[11] = 2
    r[12] = -1
    r[13] = -1
    r[14] = radius
    r[15] = proj
    return r


def emitter_from_mesh(positions, indices, world=None, radiance=(1.0, 1.0, 1.0)) -> np.ndarray:
    """The emitter record of an emissive mesh the way collectEmitterRecords derives it: the
    world-space triangle area, the geometry's bounding sphere (box centre, farthest vertex)
    moved by the world matrix and scaled by its largest axis scale."""
    pos = np.asarray(positions, np.float64).reshape(-1, 3)
    idx = np.asarray(indices, np.int64).reshape(-1, 3)
    w = np.eye(4) if world is None else np.asarray(world, np.float64)
    wp = pos @ w[:3, :3].T + w[:3, 3]
    a, b, c = wp[idx[:, 0]], wp[idx[:, 1]], wp[idx[:, 2]]
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1).sum()
    centre = (pos.min(0) + pos.max(0)) / 2
    radius = np.sqrt(((pos - centre) ** 2).sum(1).max())
    scale = np.linalg.norm(w[:3, :3], axis=0).max()
    return emitter_light(w[:3, :3] @ centre + w[:3, 3], radius * scale, area, radiance)


@dataclass
class Scene:
    bvh_nodes: np.ndarray
    tri_index: np.ndarray
    vertex_data: np.ndarray
    tri_material: np.ndarray
    materials: np.ndarray
    lights: np.ndarray
    inst_base: int
    inst_count: int
    tlas_base: int
    tlas_node_count: int
    tlas_order: np.ndarray
    meta: dict = field(default_factory=dict)
    emitter_base: int = 0
    emitter_count: int = 0
    _view: N.SceneView | None = None

    @property
    def node_count(self) -> int:
        return self.bvh_nodes.size // 8

    @property
    def tri_count(self) -> int:
        return self.tri_material.size

    def view(self) -> N.SceneView:
        if self._view is None:
            v = N.SceneView()
            v.bvh_nodes = self.bvh_nodes.ctypes.data
            v.node_count = self.node_count
            v.tri_count = self.tri_count
            v.tri_index = self.tri_index.ctypes.data
            v.vertex_data = self.vertex_data.ctypes.data
            v.vertex_count = self.vertex_data.size // 8
            v.materials_len = self.materials.size
            v.tri_material = self.tri_material.ctypes.data
            v.materials = self.materials.ctypes.data
            v.inst_base = self.inst_base
            v.inst_count = self.inst_count
            v.tlas_base = self.tlas_base
            v.tlas_node_count = self.tlas_node_count
            v.lights = self.lights.ctypes.data if self.lights.size else None
            v.light_count = self.lights.size // LIGHT_STRIDE
            v.emitter_base = self.emitter_base
            v.emitter_count = self.emitter_count
            self._view = v
        return self._view


class SceneBuilder:
    """Thin owner of an sbgi_scene_builder handle."""

    def __init__(self):
        self.L = N.lib()
        h = C.c_void_p()
        N.check(self.L.sbgi_sb_create(C.byref(h)))
        self.h = h
        self.n_instances = 0
        self.n_meshes = 0

    def __del__(self):
        if getattr(self, "h", None):
            self.L.sbgi_sb_destroy(self.h)
            self.h = None

    def _chk(self, code):
        return N.check(code, self.h, builder=True)

    def add_material(self, rec) -> int:
        rec = np.ascontiguousarray(rec, np.float32)
        i = C.c_uint32()
        self._chk(self.L.sbgi_sb_add_material(self.h, N.ptr(rec), C.byref(i)))
        return i.value

    def add_materials(self, recs) -> int:
        """n uber records [n, 28] at once; returns the id of the first."""
        recs = np.ascontiguousarray(recs, np.float32).reshape(-1, MAT_STRIDE)
        i = C.c_uint32()
        self._chk(self.L.sbgi_sb_add_materials(self.h, N.ptr(recs), recs.shape[0], C.byref(i)))
        return i.value

    def add_mesh(self, positions, indices, normals=None, uvs=None, tri_material=None) -> int:
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
        idx = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        nrm = None if normals is None else np.ascontiguousarray(normals, np.float32).reshape(-1)
        uv = None if uvs is None else np.ascontiguousarray(uvs, np.float32).reshape(-1)
        tm = None if tri_material is None else np.ascontiguousarray(tri_material, np.uint32)
        i = C.c_uint32()
        self._chk(self.L.sbgi_sb_add_mesh(self.h, N.ptr(pos), pos.size // 3, N.ptr(nrm), N.ptr(uv), N.ptr(idx),
                                          idx.size // 3, N.ptr(tm), C.byref(i)))
        self.n_meshes += 1
        return i.value

    def add_instance(self, mesh_id: int, world=None, active: bool = True) -> int:
        w = np.eye(4, dtype=np.float64) if world is None else np.asarray(world, np.float64)
        col_major = np.ascontiguousarray(w.T.reshape(-1))  # three.js Matrix4.elements
        i = C.c_uint32()
        self._chk(self.L.sbgi_sb_add_instance(self.h, mesh_id, N.ptr(col_major), int(active), C.byref(i)))
        self.n_instances += 1
        return i.value

    def add_light(self, rec):
        """One light record; emitter records (type 3) are placed after the analytic lights
        when the scene is finalized."""
        rec = np.ascontiguousarray(rec, np.float32)
        self._chk(self.L.sbgi_sb_add_light(self.h, N.ptr(rec)))

    def finalize(self, **meta) -> Scene:
        v = N.SceneView()
        self._chk(self.L.sbgi_sb_finalize(self.h, C.byref(v)))

        def grab(p, n, dt):
            if n == 0 or not p:
                return np.zeros(0, dt)
            buf = (C.c_char * (n * np.dtype(dt).itemsize)).from_address(p)
            return np.frombuffer(bytes(buf), dtype=dt).copy()

        order = np.zeros(v.inst_count, np.uint32)
        self._chk(self.L.sbgi_sb_tlas_order(self.h, N.ptr(order), v.inst_count))
        ranges = np.zeros((self.n_meshes, 6), np.uint32)
        for m in range(self.n_meshes):
            self._chk(self.L.sbgi_sb_mesh_ranges(self.h, m, N.ptr(ranges[m])))
        meta = dict(meta, mesh_ranges=ranges)
        return Scene(
            bvh_nodes=grab(v.bvh_nodes, v.node_count * 8, np.uint32),
            tri_index=grab(v.tri_index, v.tri_count * 3, np.uint32),
            vertex_data=grab(v.vertex_data, v.vertex_count * 8, np.float32),
            tri_material=grab(v.tri_material, v.tri_count, np.uint32),
            materials=grab(v.materials, v.materials_len, np.float32),
            lights=grab(v.lights, v.light_count * LIGHT_STRIDE, np.float32),
            inst_base=v.inst_base, inst_count=v.inst_count, tlas_base=v.tlas_base,
            tlas_node_count=v.tlas_node_count, tlas_order=order, meta=dict(meta),
            emitter_base=v.emitter_base, emitter_count=v.emitter_count)

    def set_transform(self, instance_id: int, world) -> int:
        w = np.ascontiguousarray(np.asarray(world, np.float64).T.reshape(-1))
        slot = C.c_uint32()
        self._chk(self.L.sbgi_sb_set_transform(self.h, instance_id, N.ptr(w), C.byref(slot)))
        return slot.value
