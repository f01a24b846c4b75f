"""Seeded procedural scenes for the probe-update workloads.

Every stream is SplitMix64 with master seed 0x5BA11000 + workload id, so a scene is
a pure function of (workload, size knobs). Geometry is generated with numpy and
assembled by the native SceneBuilder into the reference layouts.

  room_boxes()  10x10x10 m double-sided room (12 triangles) + 64 yawed boxes (12 each),
                one merged BLAS, one identity instance; 1 point light (intensity 10,
                decay 2) + 1 directional light; 16^3 probes inset 0.25 m, 64 rays.
  instanced()   8 procedural meshes (UV sphere, torus, fbm-displaced grid) instanced
                256 times with random TRS in a 40x20x25 m box, 262,144 triangles total,
                16 point lights log-uniform [1, 100]; 32x16x24 probes, 256 rays.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .scene import SceneBuilder, Scene, uber_material, point_light, directional_light

MASK64 = (1 << 64) - 1
SEED_BASE = 0x5BA11000


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def log_uniform(self, lo: float, hi: float) -> float:
        return math.exp(self.uniform(math.log(lo), math.log(hi)))

    def fork(self, tag: int) -> "SplitMix64":
        return SplitMix64(self.next_u64() ^ (tag * 0x632BE59BD9B4E019 & MASK64))


@dataclass
class Workload:
    name: str
    scene: Scene
    grid_min: tuple
    grid_size: tuple
    res: tuple
    rays: int
    depth_oct: int
    frame_jitter: float
    hysteresis: float = 0.97
    bounces: int = 1

    @property
    def probes(self) -> int:
        return self.res[0] * self.res[1] * self.res[2]


# ---------------------------------------------------------------- geometry helpers
def box_mesh(center, half, yaw):
    """12-triangle box, outward counter-clockwise winding."""
    hx, hy, hz = half
    c, s = math.cos(yaw), math.sin(yaw)
    corners = []
    for i in range(8):
        x = hx if i & 1 else -hx
        y = hy if i & 2 else -hy
        z = hz if i & 4 else -hz
        corners.append((center[0] + c * x + s * z, center[1] + y, center[2] - s * x + c * z))
    v = np.array(corners, np.float32)
    # faces as quads (a, b, c, d) counter-clockwise seen from outside
    quads = [(0, 4, 6, 2), (1, 3, 7, 5), (0, 1, 5, 4), (2, 6, 7, 3), (0, 2, 3, 1), (4, 5, 7, 6)]
    tris = []
    for a, b, cc, d in quads:
        tris += [(a, b, cc), (a, cc, d)]
    return v, np.array(tris, np.uint32)


def room_mesh(size=10.0):
    """12 wall triangles of the [0,size]^3 room (double-sided material)."""
    v, t = box_mesh((size / 2, size / 2, size / 2), (size / 2,) * 3, 0.0)
    return v, t[:, ::-1].copy()  # inward-facing


def uv_sphere(n_lat: int, n_lon: int, radius=1.0):
    verts = [(0.0, radius, 0.0)]
    for i in range(1, n_lat):
        th = math.pi * i / n_lat
        for j in range(n_lon):
            ph = 2 * math.pi * j / n_lon
            verts.append((radius * math.sin(th) * math.cos(ph), radius * math.cos(th), radius * math.sin(th) * math.sin(ph)))
    verts.append((0.0, -radius, 0.0))
    v = np.array(verts, np.float32)
    tris = []
    ring = lambda i, j: 1 + (i - 1) * n_lon + (j % n_lon)
    for j in range(n_lon):
        tris.append((0, ring(1, j + 1), ring(1, j)))
    for i in range(1, n_lat - 1):
        for j in range(n_lon):
            a, b, c, d = ring(i, j), ring(i, j + 1), ring(i + 1, j + 1), ring(i + 1, j)
            tris += [(a, b, c), (a, c, d)]
    last = len(verts) - 1
    for j in range(n_lon):
        tris.append((last, ring(n_lat - 1, j), ring(n_lat - 1, j + 1)))
    return v, np.array(tris, np.uint32)


def torus(n_u: int, n_v: int, R=1.0, r=0.35):
    u = np.arange(n_u) * (2 * math.pi / n_u)
    w = np.arange(n_v) * (2 * math.pi / n_v)
    uu, ww = np.meshgrid(u, w, indexing="ij")
    x = (R + r * np.cos(ww)) * np.cos(uu)
    y = r * np.sin(ww)
    z = (R + r * np.cos(ww)) * np.sin(uu)
    v = np.stack([x, y, z], -1).reshape(-1, 3).astype(np.float32)
    i, j = np.meshgrid(np.arange(n_u), np.arange(n_v), indexing="ij")
    a = i * n_v + j
    b = ((i + 1) % n_u) * n_v + j
    c = ((i + 1) % n_u) * n_v + (j + 1) % n_v
    d = i * n_v + (j + 1) % n_v
    t = np.concatenate([np.stack([a, d, c], -1).reshape(-1, 3), np.stack([a, c, b], -1).reshape(-1, 3)])
    return v, t.astype(np.uint32)


def _value_noise(x, y, seed: int):
    xi, yi = np.floor(x).astype(np.int64), np.floor(y).astype(np.int64)
    xf, yf = x - xi, y - yi

    def h(ix, iy):
        z = (ix * 0x27D4EB2D + iy * 0x165667B1 + seed * 0x9E3779B1) & 0xFFFFFFFF
        z = ((z ^ (z >> 15)) * 0x85EBCA6B) & 0xFFFFFFFF
        z = ((z ^ (z >> 13)) * 0xC2B2AE35) & 0xFFFFFFFF
        z = z ^ (z >> 16)
        return (z & 0xFFFFFF) / float(0xFFFFFF) * 2.0 - 1.0

    sx, sy = xf * xf * (3 - 2 * xf), yf * yf * (3 - 2 * yf)
    n0 = h(xi, yi) * (1 - sx) + h(xi + 1, yi) * sx
    n1 = h(xi, yi + 1) * (1 - sx) + h(xi + 1, yi + 1) * sx
    return n0 * (1 - sy) + n1 * sy


def fbm_grid(n: int, size=2.0, amplitude=0.5, seed=1):
    """n x n quads over [-size/2, size/2]^2 displaced in y by 4-octave fbm of amplitude 0.5 m."""
    g = np.linspace(-size / 2, size / 2, n + 1)
    xx, zz = np.meshgrid(g, g, indexing="ij")
    hgt = np.zeros_like(xx)
    amp, freq = 1.0, 1.5
    norm = 0.0
    for o in range(4):
        hgt += amp * _value_noise(xx * freq + 17.0 * o, zz * freq - 11.0 * o, seed + o)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    hgt = hgt / norm * amplitude
    v = np.stack([xx, hgt, zz], -1).reshape(-1, 3).astype(np.float32)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = i * (n + 1) + j
    b = (i + 1) * (n + 1) + j
    c = (i + 1) * (n + 1) + j + 1
    d = i * (n + 1) + j + 1
    t = np.concatenate([np.stack([a, d, c], -1).reshape(-1, 3), np.stack([a, c, b], -1).reshape(-1, 3)])
    return v, t.astype(np.uint32)


def mesh_for(kind: int, tris: float, seed: int):
    if kind == 0:
        n_lat = max(3, int(round(math.sqrt(tris / 4.0))))
        return uv_sphere(n_lat, 2 * n_lat)
    if kind == 1:
        n_v = max(3, int(round(math.sqrt(tris / 4.0))))
        return torus(2 * n_v, n_v)
    n = max(2, int(round(math.sqrt(tris / 2.0))))
    return fbm_grid(n, seed=seed)


def random_rotation(rng: SplitMix64) -> np.ndarray:
    """Uniform random rotation (Shoemake's quaternion method)."""
    u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
    q = (math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
         math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3))
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def trs(t, R, s) -> np.ndarray:
    m = np.eye(4)
    m[:3, :3] = R * s
    m[:3, 3] = t
    return m


# ---------------------------------------------------------------- workloads
def room_boxes(res=(16, 16, 16), rays: int = 64, n_boxes: int = 64, depth_oct: int = 6, seed_id: int = 1,
               hysteresis: float = 0.97) -> Workload:
    """Room + boxes: one BLAS, one identity instance, 1 point + 1 directional light."""
    rng = SplitMix64(SEED_BASE + seed_id)
    b = SceneBuilder()
    wall = b.add_material(uber_material((0.5, 0.5, 0.5), side=2))
    rv, rt = room_mesh(10.0)
    verts, tris, tmat = [rv], [rt], [np.full(len(rt), wall, np.uint32)]
    base = len(rv)
    for _ in range(n_boxes):
        half = (rng.uniform(0.3, 2.0) / 2, rng.uniform(0.3, 2.0) / 2, rng.uniform(0.3, 2.0) / 2)
        yaw = rng.uniform(0.0, 2 * math.pi)
        ex = abs(math.cos(yaw)) * half[0] + abs(math.sin(yaw)) * half[2]
        ez = abs(math.sin(yaw)) * half[0] + abs(math.cos(yaw)) * half[2]
        ctr = (rng.uniform(ex, 10 - ex), rng.uniform(half[1], 10 - half[1]), rng.uniform(ez, 10 - ez))
        alb = (rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8))
        m = b.add_material(uber_material(alb))
        bv, bt = box_mesh(ctr, half, yaw)
        verts.append(bv)
        tris.append(bt + base)
        tmat.append(np.full(len(bt), m, np.uint32))
        base += len(bv)
    mesh = b.add_mesh(np.concatenate(verts), np.concatenate(tris), tri_material=np.concatenate(tmat))
    b.add_instance(mesh)
    lp = (rng.uniform(1, 9), rng.uniform(6, 9), rng.uniform(1, 9))
    b.add_light(point_light(lp, intensity=10.0, distance=0.0, decay=2.0))
    d = np.array([rng.uniform(-1, 1), -rng.uniform(0.2, 1.0), rng.uniform(-1, 1)])
    b.add_light(directional_light((5.0, 20.0, 5.0), np.array((5.0, 20.0, 5.0)) + d, intensity=1.0))
    scene = b.finalize(workload="room_boxes")
    seed = SEED_BASE + seed_id
    jitter = (seed * 0.6180339887) % 1.0
    return Workload("room_boxes", scene, (0.25, 0.25, 0.25), (9.5, 9.5, 9.5), tuple(res), rays, depth_oct,
                    float(np.float32(jitter)), hysteresis)


def room_lights(res=(8, 8, 8), rays: int = 64, n_boxes: int = 32, n_point: int = 24, n_emitters: int = 4,
                depth_oct: int = 6, seed_id: int = 21, hysteresis: float = 0.97) -> Workload:
    """The room of room_boxes lit by ranged lights and emitters: n_point point lights with
    range U[1.5, 5] m (and two spot lights), one directional light, and n_emitters
    emissive ceiling panels that are both geometry and emitter records (type 3)."""
    from .scene import emitter_from_mesh

    rng = SplitMix64(SEED_BASE + seed_id)
    b = SceneBuilder()
    wall = b.add_material(uber_material((0.5, 0.5, 0.5), side=2))
    rv, rt = room_mesh(10.0)
    verts, tris, tmat = [rv], [rt], [np.full(len(rt), wall, np.uint32)]
    base = len(rv)
    for _ in range(n_boxes):
        half = (rng.uniform(0.3, 1.5) / 2, rng.uniform(0.3, 1.5) / 2, rng.uniform(0.3, 1.5) / 2)
        yaw = rng.uniform(0.0, 2 * math.pi)
        ctr = (rng.uniform(1.0, 9.0), rng.uniform(half[1], 7.0), rng.uniform(1.0, 9.0))
        alb = (rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8))
        m = b.add_material(uber_material(alb))
        bv, bt = box_mesh(ctr, half, yaw)
        verts.append(bv)
        tris.append(bt + base)
        tmat.append(np.full(len(bt), m, np.uint32))
        base += len(bv)
    emitters = []
    for i in range(n_emitters):
        ctr = (rng.uniform(1.5, 8.5), 9.6, rng.uniform(1.5, 8.5))
        half = (rng.uniform(0.2, 0.6), 0.02, rng.uniform(0.2, 0.6))
        radiance = (rng.uniform(2, 8), rng.uniform(2, 8), rng.uniform(2, 8))
        m = b.add_material(uber_material((0.8, 0.8, 0.8), emissive=radiance))
        bv, bt = box_mesh(ctr, half, 0.0)
        verts.append(bv)
        tris.append(bt + base)
        tmat.append(np.full(len(bt), m, np.uint32))
        base += len(bv)
        emitters.append(emitter_from_mesh(bv, bt, radiance=radiance))
    mesh = b.add_mesh(np.concatenate(verts), np.concatenate(tris), tri_material=np.concatenate(tmat))
    b.add_instance(mesh)
    # emitter records first: the builder still places them after the analytic lights
    for rec in emitters:
        b.add_light(rec)
    for i in range(n_point):
        lp = (rng.uniform(0.5, 9.5), rng.uniform(0.5, 9.5), rng.uniform(0.5, 9.5))
        col = (rng.uniform(0.5, 1.0), rng.uniform(0.5, 1.0), rng.uniform(0.5, 1.0))
        b.add_light(point_light(lp, col, intensity=rng.log_uniform(2.0, 20.0), distance=rng.uniform(1.5, 5.0),
                                decay=2.0))
    for i in range(2):
        lp = np.array((rng.uniform(2, 8), 9.0, rng.uniform(2, 8)))
        r = point_light(lp, intensity=15.0, distance=8.0, decay=1.0)
        r[0] = 2  # spot, pointing down
        r[4:7] = (0.0, -1.0, 0.0)
        r[12] = math.cos(0.6)
        r[13] = math.cos(0.6 * 0.7)
        b.add_light(r)
    d = np.array([rng.uniform(-1, 1), -rng.uniform(0.2, 1.0), rng.uniform(-1, 1)])
    b.add_light(directional_light((5.0, 20.0, 5.0), np.array((5.0, 20.0, 5.0)) + d, intensity=0.5))
    scene = b.finalize(workload="room_lights")
    seed = SEED_BASE + seed_id
    jitter = (seed * 0.6180339887) % 1.0
    return Workload("room_lights", scene, (0.25, 0.25, 0.25), (9.5, 9.5, 9.5), tuple(res), rays, depth_oct,
                    float(np.float32(jitter)), hysteresis)


def instanced(res=(32, 16, 24), rays: int = 256, total_tris: int = 262144, n_instances: int = 256,
              n_lights: int = 16, depth_oct: int = 14, seed_id: int = 2, hysteresis: float = 0.97,
              light_range: float = 0.0, n_emitters: int = 0) -> Workload:
    """8 procedural meshes x 256 random-TRS instances in a 40x20x25 m box, 16 point lights."""
    rng = SplitMix64(SEED_BASE + seed_id)
    box = np.array([40.0, 20.0, 25.0])
    n_mesh = 8
    target = [rng.log_uniform(1000, 64000) for _ in range(n_mesh)]
    assign = [int(rng.uniform(0, n_mesh)) % n_mesh for _ in range(n_instances)]
    counts = np.bincount(assign, minlength=n_mesh)
    scale = total_tris / float(np.dot(counts, target))
    # quantized mesh resolutions: rescale until the instanced total is within +-1% of the target
    geo = None
    for _ in range(8):
        geo = [mesh_for(m % 3, target[m] * scale, seed=m + 7) for m in range(n_mesh)]
        actual = int(np.dot(counts, [len(t) for _, t in geo]))
        if abs(actual - total_tris) <= 0.01 * total_tris:
            break
        scale *= total_tris / float(actual)
    b = SceneBuilder()
    meshes, mesh_tris = [], []
    for m in range(n_mesh):
        v, t = geo[m]
        alb = (rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8))
        mat = b.add_material(uber_material(alb, side=2 if m % 3 == 2 else 0))
        meshes.append(b.add_mesh(v, t, tri_material=np.full(len(t), mat, np.uint32)))
        mesh_tris.append(len(t))
    actual = int(np.dot(counts, mesh_tris))
    for i in range(n_instances):
        R = random_rotation(rng)
        s = rng.uniform(0.5, 2.0)
        t = (rng.uniform(0, box[0]), rng.uniform(0, box[1]), rng.uniform(0, box[2]))
        b.add_instance(meshes[assign[i]], trs(t, R, s))
    for _ in range(n_lights):
        p = (rng.uniform(0, box[0]), rng.uniform(0, box[1]), rng.uniform(0, box[2]))
        b.add_light(point_light(p, intensity=rng.log_uniform(1.0, 100.0), distance=light_range, decay=2.0))
    if n_emitters:
        from .scene import emitter_light
        for _ in range(n_emitters):  # emissive spheres, given as records only
            c = (rng.uniform(0, box[0]), rng.uniform(0, box[1]), rng.uniform(0, box[2]))
            r = rng.uniform(0.3, 1.0)
            b.add_light(emitter_light(c, r, 4 * math.pi * r * r, (rng.uniform(2, 8), rng.uniform(2, 8),
                                                                   rng.uniform(2, 8))))
    scene = b.finalize(workload="instanced", instanced_triangles=actual, unique_triangles=int(sum(mesh_tris)))
    seed = SEED_BASE + seed_id
    jitter = (seed * 0.6180339887) % 1.0
    return Workload("instanced", scene, (0.0, 0.0, 0.0), tuple(box.tolist()), tuple(res), rays, depth_oct,
                    float(np.float32(jitter)), hysteresis, bounces=2)


def differentiable(res=(16, 16, 16), rays: int = 128, n_mesh: int = 4, n_instances: int = 64, torus_n: int = 16,
                   n_lights: int = 16, box: float = 20.0, seed_id: int = 4):
    """Differentiable workload: n_mesh tori of 4 * torus_n^2 triangles (1024 at the
    default), n_instances random-TRS instances in a box, n_lights point lights with
    log-uniform intensity [1, 100], per-instance-triangle albedo U[0.1, 0.9] and a
    target albedo perturbed by U[-0.2, 0.2] (multiplicative, next seed).

    Returns (workload, albedo (n, 3) float32, target_albedo (n, 3) float32)."""
    rng = SplitMix64(SEED_BASE + seed_id)
    b = SceneBuilder()
    mat = b.add_material(uber_material((0.5, 0.5, 0.5)))
    meshes = []
    for m in range(n_mesh):
        R_ = 1.0
        r_ = 0.2 + 0.15 * m
        v, t = torus(2 * torus_n, torus_n, R=R_, r=r_)
        meshes.append(b.add_mesh(v, t, tri_material=np.full(len(t), mat, np.uint32)))
    for i in range(n_instances):
        R = random_rotation(rng)
        s = rng.uniform(0.75, 2.0)
        tr = (rng.uniform(0, box), rng.uniform(0, box), rng.uniform(0, box))
        b.add_instance(meshes[i % n_mesh], trs(tr, R, s))
    for _ in range(n_lights):
        p = (rng.uniform(0, box), rng.uniform(0, box), rng.uniform(0, box))
        b.add_light(point_light(p, intensity=rng.log_uniform(1.0, 100.0), distance=0.0, decay=2.0))
    scene = b.finalize(workload="differentiable")
    n_alb = n_instances * 4 * torus_n * torus_n
    gen = np.random.default_rng(SEED_BASE + seed_id)
    albedo = gen.uniform(0.1, 0.9, (n_alb, 3)).astype(np.float32)
    gen2 = np.random.default_rng(SEED_BASE + seed_id + 1)
    target = (albedo * (1.0 + gen2.uniform(-0.2, 0.2, albedo.shape))).astype(np.float32)
    seed = SEED_BASE + seed_id
    jitter = float(np.float32((seed * 0.6180339887) % 1.0))
    w = Workload("differentiable", scene, (0.0, 0.0, 0.0), (box, box, box), tuple(res), rays, 6, jitter, 0.0, bounces=2)
    return w, albedo, target


# ---------------------------------------------------------------- dynamic workload
def tube(n_around: int, n_along: int, length: float, radius: float):
    """Open cylinder along +y from 0 to `length`: (n_along + 1) rings of n_around vertices,
    radial normals, 2 * n_around * n_along triangles wound outward."""
    ph = np.arange(n_around) * (2 * math.pi / n_around)
    y = np.linspace(0.0, length, n_along + 1)
    yy, pp = np.meshgrid(y, ph, indexing="ij")
    nrm = np.stack([np.cos(pp), np.zeros_like(pp), np.sin(pp)], -1).reshape(-1, 3)
    v = np.stack([radius * np.cos(pp), yy, radius * np.sin(pp)], -1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(n_along), np.arange(n_around), indexing="ij")
    a = i * n_around + j
    b = i * n_around + (j + 1) % n_around
    c = (i + 1) * n_around + (j + 1) % n_around
    d = (i + 1) * n_around + j
    t = np.concatenate([np.stack([a, c, b], -1).reshape(-1, 3), np.stack([a, d, c], -1).reshape(-1, 3)])
    return v.astype(np.float32), t.astype(np.uint32), nrm.astype(np.float32)


def axis_angle(axis, angle: float) -> np.ndarray:
    x, y, z = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    c, s = math.cos(angle), math.sin(angle)
    C1 = 1 - c
    return np.array([[c + x * x * C1, x * y * C1 - z * s, x * z * C1 + y * s],
                     [y * x * C1 + z * s, c + y * y * C1, y * z * C1 - x * s],
                     [z * x * C1 - y * s, z * y * C1 + x * s, c + z * z * C1]])


@dataclass
class DynamicWorkload:
    """Skinned tubes + rigid instances. skin_sets / n_bones feed ProbeField.set_skins;
    bones(t) is the palette at time t (s); rigid_step() advances the random walk one frame
    and returns (TLAS slots, world matrices)."""
    name: str
    scene: Scene
    cascades: list
    rays: int
    depth_oct: int
    frame_jitter: float
    hysteresis: float
    skin_sets: list
    n_bones: int
    tube_len: np.ndarray
    tube_axis: np.ndarray
    bone_phase: np.ndarray
    rigid_slots: np.ndarray
    slot_worlds: np.ndarray
    rigid_pos: np.ndarray
    rigid_rot: np.ndarray
    rigid_scale: np.ndarray
    box: np.ndarray
    walk_rng: np.random.Generator
    amplitude: float = math.radians(30.0)
    frequency: float = 0.5
    fps: float = 60.0
    bounces: int = 2

    @property
    def probes(self) -> int:
        return sum(r[0] * r[1] * r[2] for _, _, r in self.cascades)

    def bones(self, t: float) -> np.ndarray:
        """[n_bones, 3, 4]: bone b of tube k swings about the tube's axis by
        amplitude * sin(2 pi f t + phase) around its pivot at height b * L / 4, chained."""
        out = np.zeros((self.n_bones, 3, 4), np.float64)
        for k in range(len(self.tube_len)):
            M = np.eye(4)
            seg = self.tube_len[k] / 4.0
            for b in range(4):
                th = self.amplitude * math.sin(2 * math.pi * self.frequency * t + self.bone_phase[k, b])
                L = np.eye(4)
                L[:3, :3] = axis_angle(self.tube_axis[k], th)
                piv = np.array([0.0, b * seg, 0.0])
                L[:3, 3] = piv - L[:3, :3] @ piv
                M = M @ L
                out[k * 4 + b] = M[:3, :]
        return out.astype(np.float32)

    def rigid_step(self, sigma: float = 0.05, sigma_rot: float = math.radians(1.0)):
        g = self.walk_rng
        n = len(self.rigid_slots)
        self.rigid_pos += g.normal(0.0, sigma, (n, 3))
        lo, hi = np.zeros(3), self.box
        self.rigid_pos = np.where(self.rigid_pos < lo, 2 * lo - self.rigid_pos, self.rigid_pos)
        self.rigid_pos = np.where(self.rigid_pos > hi, 2 * hi - self.rigid_pos, self.rigid_pos)
        worlds = []
        for i in range(n):
            ax = g.normal(size=3)
            self.rigid_rot[i] = axis_angle(ax, g.normal(0.0, sigma_rot)) @ self.rigid_rot[i]
            worlds.append(trs(self.rigid_pos[i], self.rigid_rot[i], self.rigid_scale[i]))
        return self.rigid_slots, worlds


def dynamic(res=(24, 24, 24), n_cascades: int = 4, rays: int = 128, n_tubes: int = 32, tube_seg=(128, 128),
            n_rigid: int = 128, n_lights: int = 16, depth_oct: int = 14, seed_id: int = 5,
            hysteresis: float = 0.97, box=(40.0, 20.0, 25.0)) -> DynamicWorkload:
    """32 skinned tubes (4-bone chains, 128 x 128 segments = 32,768 triangles each, 1,048,576
    in all) standing in a 40x20x25 m box with 128 rigid instances of a sphere, a torus and a
    box doing a seeded random walk; 16 point lights log-uniform [1, 100]; 4 cascades of
    24^3 probes centred on the box, each twice the extent of the previous (the largest is
    the box), 128 rays."""
    rng = SplitMix64(SEED_BASE + seed_id)
    box = np.array(box, np.float64)
    b = SceneBuilder()
    tube_mat = b.add_material(uber_material((0.7, 0.6, 0.5), side=2))
    tubes, lens, axes = [], [], []
    for k in range(n_tubes):
        L = rng.uniform(4.0, 10.0)
        r = rng.uniform(0.2, 0.6)
        v, t, nrm = tube(tube_seg[0], tube_seg[1], L, r)
        tubes.append((b.add_mesh(v, t, normals=nrm, tri_material=np.full(len(t), tube_mat, np.uint32)), v))
        lens.append(L)
        a = rng.uniform(0, 2 * math.pi)
        axes.append((math.cos(a), 0.0, math.sin(a)))
    rigid_meshes = []
    for kind, (v, t) in enumerate([uv_sphere(16, 32, 0.6), torus(32, 16, 0.6, 0.2), box_mesh((0, 0, 0), (0.5,) * 3, 0)]):
        m = b.add_material(uber_material((rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8))))
        rigid_meshes.append(b.add_mesh(v, t, tri_material=np.full(len(t), m, np.uint32)))
    worlds = []
    for k in range(n_tubes):
        yaw = rng.uniform(0, 2 * math.pi)
        R = axis_angle((0, 1, 0), yaw) @ axis_angle((1, 0, 0), rng.uniform(-0.2, 0.2))
        pos = (rng.uniform(0.1, 0.9) * box[0], rng.uniform(0, 0.1) * box[1], rng.uniform(0.1, 0.9) * box[2])
        worlds.append(trs(pos, R, 1.0))
        b.add_instance(tubes[k][0], worlds[-1])
    rigid_ids, rpos, rrot, rscale = [], [], [], []
    for i in range(n_rigid):
        R = random_rotation(rng)
        s = rng.uniform(0.5, 2.0)
        p = np.array([rng.uniform(0, box[0]), rng.uniform(0, box[1]), rng.uniform(0, box[2])])
        worlds.append(trs(p, R, s))
        rigid_ids.append(b.add_instance(rigid_meshes[i % 3], worlds[-1]))
        rpos.append(p)
        rrot.append(R)
        rscale.append(s)
    for _ in range(n_lights):
        p = (rng.uniform(0, box[0]), rng.uniform(0, box[1]), rng.uniform(0, box[2]))
        b.add_light(point_light(p, intensity=rng.log_uniform(1.0, 100.0), distance=0.0, decay=2.0))
    scene = b.finalize(workload="dynamic", skinned_triangles=n_tubes * 2 * tube_seg[0] * tube_seg[1])
    ranges = scene.meta["mesh_ranges"]
    slot_of = np.argsort(scene.tlas_order)  # source instance id -> TLAS slot
    # 4-bone weights: Gaussian in height around each bone's centre, normalised
    sets = []
    for k in range(n_tubes):
        mid, v = tubes[k]
        vs, vc, _, _, br, be = (int(x) for x in ranges[mid])
        seg = lens[k] / 4.0
        c = (np.arange(4) + 0.5) * seg
        w = np.exp(-(((v[:, 1:2] - c[None, :]) / (0.6 * seg)) ** 2))
        w = (w / w.sum(1, keepdims=True)).astype(np.float32)
        j = (k * 4 + np.arange(4, dtype=np.uint16))[None, :].repeat(vc, 0)
        sets.append((vs, vc, br, be, j, w))
    gen = np.random.default_rng(SEED_BASE + seed_id)
    phase = gen.uniform(0, 2 * math.pi, (n_tubes, 4))
    centre = box / 2
    cascades = []
    for c in range(n_cascades):
        size = box / (2 ** (n_cascades - 1 - c))
        cascades.append((tuple((centre - size / 2).tolist()), tuple(size.tolist()), tuple(res)))
    seed = SEED_BASE + seed_id
    jitter = float(np.float32((seed * 0.6180339887) % 1.0))
    return DynamicWorkload("dynamic", scene, cascades, rays, depth_oct, jitter, hysteresis, sets, n_tubes * 4,
                           np.array(lens), np.array(axes), phase, slot_of[np.array(rigid_ids)].astype(np.uint32),
                           np.array(worlds)[scene.tlas_order], np.array(rpos), np.array(rrot), np.array(rscale), box,
                           np.random.default_rng(SEED_BASE + seed_id + 1))


# ---------------------------------------------------------------- batched workload
def boxes_vectorized(ctr, half, yaw):
    """n yawed boxes at once: vertices (8n, 3) and outward triangles (12n, 3)."""
    n = len(ctr)
    sx = np.array([-1, 1, -1, 1, -1, 1, -1, 1], np.float64)
    sy = np.array([-1, -1, 1, 1, -1, -1, 1, 1], np.float64)
    sz = np.array([-1, -1, -1, -1, 1, 1, 1, 1], np.float64)
    x = half[:, 0:1] * sx[None]
    y = half[:, 1:2] * sy[None]
    z = half[:, 2:3] * sz[None]
    c, s = np.cos(yaw)[:, None], np.sin(yaw)[:, None]
    v = np.stack([ctr[:, 0:1] + c * x + s * z, ctr[:, 1:2] + y, ctr[:, 2:3] - s * x + c * z], -1)
    quads = np.array([(0, 4, 6, 2), (1, 3, 7, 5), (0, 1, 5, 4), (2, 6, 7, 3), (0, 2, 3, 1), (4, 5, 7, 6)])
    tri = np.concatenate([quads[:, [0, 1, 2]], quads[:, [0, 2, 3]]], 0)  # 12 x 3 per box
    t = tri[None, :, :] + (np.arange(n) * 8)[:, None, None]
    return v.reshape(-1, 3).astype(np.float32), t.reshape(-1, 3).astype(np.uint32)


def room_boxes_scaled(n_tris: int, res, rays: int, n_cascades: int = 1, seed_id: int = 100, depth_oct: int = 6,
                      hysteresis: float = 0.97):
    """The room-and-boxes recipe grown to about n_tris triangles: (n_tris - 12) / 12 boxes
    (sizes U[0.3, 2] m, random yaw, albedo U[0.2, 0.8]^3 each) in a cubic room whose side
    keeps 64 boxes per 1000 m^3, 1 point light (intensity 10) + 1 directional light;
    n_cascades grids of res probes centred in the room, each twice the extent of the
    previous, the largest inset 2.5 % from the walls. Returns (scene, cascades, jitter)."""
    rng = np.random.default_rng(SEED_BASE + seed_id)
    n_boxes = max(1, int(round((n_tris - 12) / 12)))
    L = 10.0 * (n_boxes / 64.0) ** (1.0 / 3.0)
    half = rng.uniform(0.3, 2.0, (n_boxes, 3)) / 2
    yaw = rng.uniform(0.0, 2 * math.pi, n_boxes)
    ex = np.abs(np.cos(yaw)) * half[:, 0] + np.abs(np.sin(yaw)) * half[:, 2]
    ez = np.abs(np.sin(yaw)) * half[:, 0] + np.abs(np.cos(yaw)) * half[:, 2]
    ctr = np.stack([rng.uniform(ex, L - ex), rng.uniform(half[:, 1], L - half[:, 1]), rng.uniform(ez, L - ez)], 1)
    alb = rng.uniform(0.2, 0.8, (n_boxes, 3))
    b = SceneBuilder()
    wall = b.add_material(uber_material((0.5, 0.5, 0.5), side=2))
    recs = np.tile(uber_material(), (n_boxes, 1))
    recs[:, 0:3] = alb
    first = b.add_materials(recs)
    rv, rt = room_mesh(L)
    bv, bt = boxes_vectorized(ctr, half, yaw)
    v = np.concatenate([rv, bv])
    t = np.concatenate([rt, bt + len(rv)])
    tm = np.concatenate([np.full(len(rt), wall, np.uint32), (first + np.repeat(np.arange(n_boxes), 12)).astype(np.uint32)])
    b.add_instance(b.add_mesh(v, t, tri_material=tm))
    lp = (rng.uniform(0.1, 0.9) * L, rng.uniform(0.6, 0.9) * L, rng.uniform(0.1, 0.9) * L)
    b.add_light(point_light(lp, intensity=10.0, distance=0.0, decay=2.0))
    d = np.array([rng.uniform(-1, 1), -rng.uniform(0.2, 1.0), rng.uniform(-1, 1)])
    top = np.array((L / 2, 2 * L, L / 2))
    b.add_light(directional_light(tuple(top), top + d, intensity=1.0))
    scene = b.finalize(workload="room_boxes_scaled", boxes=n_boxes, room=L)
    centre = np.full(3, L / 2)
    big = np.full(3, 0.95 * L)
    cascades = []
    for k in range(n_cascades):
        size = big / (2 ** (n_cascades - 1 - k))
        cascades.append((tuple((centre - size / 2).tolist()), tuple(size.tolist()), tuple(res)))
    jitter = float(np.float32(((SEED_BASE + seed_id) * 0.6180339887) % 1.0))
    return scene, cascades, jitter


def batch_specs(n_fields: int = 48, seed_id: int = 3):
    """Per field of the batched workload: (triangles log-uniform [2k, 2M], probes per cascade
    log-uniform [512, 32768] as a cubic grid, rays from {32, 64, 128, 256}, cascades 1-4)."""
    rng = SplitMix64(SEED_BASE + seed_id)
    out = []
    for i in range(n_fields):
        tris = int(round(rng.log_uniform(2000, 2_000_000)))
        probes = rng.log_uniform(512, 32768)
        r = max(8, min(32, int(round(probes ** (1.0 / 3.0)))))
        rays = (32, 64, 128, 256)[int(rng.uniform(0, 4)) % 4]
        casc = 1 + int(rng.uniform(0, 4)) % 4
        out.append(dict(n_tris=tris, res=(r, r, r), rays=rays, n_cascades=casc, seed_id=1000 + i))
    return out
