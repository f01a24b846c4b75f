"""Seeded random scenes against the FP32 restatement (run on an MI355X with -m gpu).

Each seed builds its own two-level scene: 3-6 meshes of the three procedural kinds (UV
sphere, torus, fBm height grid) with single-sided front, back or double-sided materials,
20-60 instances with random rotations and scales, some of them mirrored (negative
determinant: the walk's det-sign flip) and some inactive, and 2-9 lights mixing point,
ranged point, spot and directional lights. On a small probe field the GPU's hit ids, light
visibility masks and radiance are compared with the FP32 restatement in both traversal
modes, at the bars of the fixed workloads: hit ids identical except near ties (two
nearest accepted hits within 1e-6 relative), at most 1 shadow ray in 10,000 on the other
side of a graze, radiance p99.9 within 1e-4 relative on the rays with the same hit.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from speedball_amd import _native as N
from speedball_amd import scenegen as sg
from speedball_amd.field import ProbeField
from speedball_amd.scene import SceneBuilder, directional_light, point_light, spot_light, uber_material
from tests.parity import classify_mismatches, rel_err, summary

pytestmark = pytest.mark.gpu

SEEDS = [11, 23, 37, 41, 59, 73]


def random_workload(seed):
    rng = np.random.default_rng(seed)
    box = np.array([14.0, 8.0, 10.0])
    b = SceneBuilder()
    meshes = []
    for m in range(int(rng.integers(3, 7))):
        kind = m % 3
        v, t = sg.mesh_for(kind, float(rng.uniform(150, 1500)), seed=seed * 10 + m)
        col = tuple(float(x) for x in rng.uniform(0.2, 0.9, 3))
        mat = b.add_material(uber_material(col, side=int(rng.integers(0, 3))))
        meshes.append(b.add_mesh(v, t, tri_material=np.full(len(t), mat, np.uint32)))
    n_inst = int(rng.integers(20, 61))
    for i in range(n_inst):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        a, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - a * z), 2 * (x * z + a * y)],
                      [2 * (x * y + a * z), 1 - 2 * (x * x + z * z), 2 * (y * z - a * x)],
                      [2 * (x * z - a * y), 2 * (y * z + a * x), 1 - 2 * (x * x + y * y)]])
        s = float(rng.uniform(0.4, 1.6))
        M = np.eye(4)
        M[:3, :3] = R * s
        if rng.uniform() < 0.25:
            M[:3, 0] *= -1.0  # mirrored instance
        M[:3, 3] = rng.uniform(0, 1, 3) * box
        b.add_instance(meshes[int(rng.integers(0, len(meshes)))], M, active=bool(rng.uniform() > 0.1))
    for li in range(int(rng.integers(2, 10))):
        p = tuple(float(x) for x in rng.uniform(0, 1, 3) * box)
        col = tuple(float(x) for x in rng.uniform(0.3, 1.0, 3))
        kind = li % 4
        if kind == 0:
            rec = point_light(p, color=col, intensity=float(rng.uniform(5, 40)))
        elif kind == 1:
            rec = point_light(p, color=col, intensity=float(rng.uniform(5, 40)), distance=float(rng.uniform(3, 8)),
                              decay=1.5)
        elif kind == 2:
            tgt = tuple(float(x) for x in rng.uniform(0, 1, 3) * box)
            rec = spot_light(p, tgt, color=col, intensity=float(rng.uniform(10, 60)), angle=0.6, penumbra=0.3)
        else:
            tgt = tuple(float(x) for x in rng.uniform(0, 1, 3) * box)
            rec = directional_light(p, tgt, color=col, intensity=float(rng.uniform(0.5, 2.0)))
        b.add_light(rec)
    scene = b.finalize(workload=f"random_{seed}")
    jitter = float(np.float32(rng.uniform()))
    return sg.Workload(f"random_{seed}", scene, (0.0, 0.0, 0.0), tuple(box.tolist()), (8, 5, 6), 64, 6, jitter)


def _popcount(x):
    return int(np.unpackbits(np.ascontiguousarray(x).view(np.uint8)).sum())


@pytest.mark.parametrize("traversal", [N.TRAV_MT_COMPAT, N.TRAV_MT_ORDERED], ids=["compat", "ordered"])
@pytest.mark.parametrize("seed", SEEDS)
def test_random_scene_matches_oracle(seed, traversal):
    w = random_workload(seed)
    print(f"case: random scene seed {seed}, traversal {traversal}, {w.scene.tri_count} triangles, "
          f"{w.scene.inst_count} instances, {len(w.scene.lights) // 17} lights", flush=True)
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=w.depth_oct, keep_hit_ids=True,
                   traversal=traversal)
    try:
        f.set_scene(w.scene)
        f.set_counting(True)
        f.tick(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
        g_rays, g_hits = f.read_rays(), f.read_hit_ids()
        g_vis = f.read(N.BUF_SHADOW_VIS, np.uint64).reshape(-1, w.rays)
        shadow_rays = int(f.stats().shadow_rays)
    finally:
        f.close()
    o = oracle.OracleField(w.scene, w.grid_min, w.grid_size, w.res, w.rays, w.depth_oct, prec="32")
    o.trace(o.params(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis), vis=True)
    r_rays, r_hits, r_vis = o.rays_view(), o.hits_view(), o.vis.reshape(-1, w.rays)
    mism = np.any(g_hits != r_hits, axis=-1)
    slots, ks = np.nonzero(mism)
    ties, _ = classify_mismatches(oracle, w.scene, w, slots, ks, g_hits[slots, ks], r_hits[slots, ks],
                                  g_rays[slots, ks, 3], edge=0.0)
    same = ~mism
    hit = same & (r_hits[..., 0] >= 0)
    diff = _popcount((g_vis ^ r_vis)[same])
    e = rel_err(g_rays[..., :3][same], r_rays[..., :3][same])
    print(f"rays {mism.size}, hits {int(hit.sum())}, id mismatches {int(mism.sum())} (near ties {int(ties.sum())}), "
          f"shadow rays {shadow_rays}, visibility bits that differ {diff}; radiance {summary(e)}")
    assert hit.sum() > mism.size // 20  # the scene is hit
    assert int((~ties).sum()) == 0
    assert diff <= max(1, shadow_rays // 10000)
    assert rel_err(g_rays[..., 3][hit], r_rays[..., 3][hit], 0).max() <= 1e-5
    assert np.percentile(e, 99.9) <= 1e-4
