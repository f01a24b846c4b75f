"""Dynamic scene edits on the GPU against the CPU oracle: instance transforms, BLAS/TLAS
refits, device skinning and the per-frame SAH TLAS (run on an MI355X with -m gpu).

Bars: bit-exact for instance world bounds, the SAH TLAS (against the oracle's
restatement of the device build), BLAS and reference-TLAS refits, skinned vertex
positions and normals and the repacked triangles (same float operation order); after the
edits, hit triangle/instance ids identical except near ties (two nearest hits within 1e-6
relative, counted) and hit t within 1e-5 relative, traced by the FP32 oracle on the edited
scene.
"""
from __future__ import annotations

import dataclasses
import math
import time
from types import SimpleNamespace

import numpy as np
import pytest

import oracle
from speedball_amd import _native as N
from speedball_amd import scenegen as sg
from speedball_amd.field import ProbeField
from tests.parity import classify_mismatches, rel_err

pytestmark = pytest.mark.gpu

MODES = [N.TRAV_MT_COMPAT, N.TRAV_MT_ORDERED]
IDS = ["compat", "ordered"]


def scene_with(s, **repl):
    kw = {f.name: getattr(s, f.name) for f in dataclasses.fields(s) if f.name != "_view"}
    kw.update(repl)
    return sg.Scene(**kw)


def make_field(scene, casc, rays, depth_oct, traversal):
    f = ProbeField(casc[0][2], casc[0][0], casc[0][1], rays=rays, depth_oct=depth_oct, keep_hit_ids=True,
                   cascades=casc, traversal=traversal)
    f.set_scene(scene)
    return f


def inflate(b):
    """The SAH build's input from read-back world bounds (float32, the stated formula)."""
    b = b.astype(np.float32)
    mn, mx = b[:, :3], b[:, 3:]
    pad = np.float32(1e-5) * (np.abs(mx - mn) + np.abs(mn) + np.abs(mx)) + np.float32(1e-6)
    return np.concatenate([mn - pad, mx + pad], 1).astype(np.float32)


def oracle_scene_after_edits(f, base, vertex_data, nodes, worlds):
    """The edited scene as the oracle traces it: device vertex data / nodes / records, and
    the reference TLAS refit over oracle instance bounds."""
    mats = f.read(N.BUF_MATERIALS, np.float32).copy()
    b, _ = oracle.inst_bounds(mats, base.inst_base, base.inst_count, nodes, worlds)
    tl = mats[base.tlas_base:base.tlas_base + base.tlas_node_count * 12].copy()
    assert oracle.refit_tlas(tl, b.reshape(-1).copy())
    mats[base.tlas_base:base.tlas_base + base.tlas_node_count * 12] = tl
    return scene_with(base, vertex_data=np.ascontiguousarray(vertex_data.reshape(-1)),
                      bvh_nodes=np.ascontiguousarray(nodes.reshape(-1)), materials=mats), b


def check_hits(f, scene, casc, rays, depth_oct, jitter, hysteresis, cascade=0, label=""):
    gm, gs, res = casc[cascade]
    g_rays, g_hits = f.read_rays(), f.read_hit_ids()
    o = oracle.OracleField(scene, gm, gs, res, rays, depth_oct, prec="32")
    p = o.params(frame_jitter=jitter, hysteresis=hysteresis)
    o.trace(p)
    r_rays, r_hits = o.rays_view().copy(), o.hits_view().copy()
    mism = np.any(g_hits != r_hits, axis=-1)
    slots, ks = np.nonzero(mism)
    w = SimpleNamespace(res=res, grid_min=gm, grid_size=gs, rays=rays, frame_jitter=jitter)
    ties, grazes = classify_mismatches(oracle, scene, w, slots, ks, g_hits[slots, ks], r_hits[slots, ks],
                                       g_rays[slots, ks, 3], edge=0.0)
    hit = r_hits[..., 0] >= 0
    print(f"\n{label}: rays {mism.size} hits {int(hit.sum())} id mismatches {int(mism.sum())} "
          f"near ties {int(ties.sum())} edge grazes {int(grazes.sum())}", flush=True)
    assert int((~(ties | grazes)).sum()) == 0
    same = ~mism & hit
    assert rel_err(g_rays[..., 3][same], r_rays[..., 3][same], 0).max() <= 1e-5
    assert hit.sum() >= 0.02 * hit.size  # the edited scene is actually in view


def slot_worlds_from_rows(scene):
    """Exact world matrices for every slot: float64 inverse of the stored inverse rows."""
    rows = scene.materials[scene.inst_base:scene.inst_base + scene.inst_count * 28].reshape(-1, 28)[:, :12]
    out = []
    for r in rows.astype(np.float64):
        m = np.eye(4)
        m[:3, :] = r.reshape(3, 4)
        out.append(np.linalg.inv(m))
    return out


# ---------------------------------------------------------------- set_scene + transforms
def test_set_scene_sah_tlas_matches_oracle(instanced_small):
    w = instanced_small
    f = make_field(w.scene, [(w.grid_min, w.grid_size, w.res)], w.rays, w.depth_oct, N.TRAV_MT_ORDERED)
    try:
        print("case: set_scene SAH TLAS", flush=True)
        b = f.read(N.BUF_INST_BOUNDS, np.float32).reshape(-1, 6)
        tl = f.read(N.BUF_TLAS_SAH, np.uint32).reshape(-1, 8)
        assert np.array_equal(tl, oracle.build_sah_tlas(inflate(b)))
    finally:
        f.close()


@pytest.mark.parametrize("traversal", MODES, ids=IDS)
def test_instance_transforms_match_oracle(instanced_small, traversal):
    """Every slot gets an exact world matrix, a third of them move; bounds, TLAS and the
    traced hits follow."""
    w = instanced_small
    s = w.scene
    f = make_field(s, [(w.grid_min, w.grid_size, w.res)], w.rays, w.depth_oct, traversal)
    try:
        worlds = slot_worlds_from_rows(s)
        rng = np.random.default_rng(31)
        moved = rng.choice(s.inst_count, s.inst_count // 3, replace=False)
        for i in moved:
            d = np.eye(4)
            d[:3, :3] = sg.axis_angle(rng.normal(size=3), rng.uniform(-0.5, 0.5))
            d[:3, 3] = rng.normal(0, 1.5, 3)
            worlds[i] = d @ worlds[i]
        print(f"case: transforms {IDS[traversal]}", flush=True)
        f.set_instance_transforms(np.arange(s.inst_count), worlds)
        wd = np.stack([m.T.reshape(-1) for m in worlds])
        nodes = f.read(N.BUF_BVH_NODES, np.uint32)
        mats = f.read(N.BUF_MATERIALS, np.float32)
        b64, pad = oracle.inst_bounds(mats, s.inst_base, s.inst_count, nodes, wd)
        assert np.array_equal(f.read(N.BUF_INST_BOUNDS, np.float32).reshape(-1, 6), b64)
        if traversal == N.TRAV_MT_ORDERED:
            assert np.array_equal(f.read(N.BUF_TLAS_SAH, np.uint32).reshape(-1, 8), oracle.build_sah_tlas(pad))
        else:
            tl = s.materials[s.tlas_base:s.tlas_base + s.tlas_node_count * 12].copy()
            assert oracle.refit_tlas(tl, b64.reshape(-1).copy())
            assert np.array_equal(mats[s.tlas_base:s.tlas_base + s.tlas_node_count * 12], tl)
        f.tick(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
        scene2, _ = oracle_scene_after_edits(f, s, s.vertex_data, nodes, wd)
        check_hits(f, scene2, [(w.grid_min, w.grid_size, w.res)], w.rays, w.depth_oct, w.frame_jitter,
                   w.hysteresis, label=f"transforms {IDS[traversal]}")
    finally:
        f.close()


@pytest.mark.parametrize("traversal", MODES, ids=IDS)
def test_vertex_update_and_refit_match_oracle(instanced_small, traversal):
    """updateDeforms path: new vertex data through update_ranges, then the device refit of
    every instanced BLAS range and the TLAS."""
    w = instanced_small
    s = w.scene
    f = make_field(s, [(w.grid_min, w.grid_size, w.res)], w.rays, w.depth_oct, traversal)
    try:
        worlds = slot_worlds_from_rows(s)
        f.set_instance_transforms(np.arange(s.inst_count), worlds)
        rng = np.random.default_rng(32)
        vd = s.vertex_data.copy().reshape(-1, 8)
        vd[:, :3] *= np.float32(1.0) + rng.uniform(-0.08, 0.08, (vd.shape[0], 1)).astype(np.float32)
        vd = np.ascontiguousarray(vd.reshape(-1))
        print(f"case: refit {IDS[traversal]}", flush=True)
        f.update_ranges(N.BUF_VERTEX_DATA, [0, vd.size], vd)
        inst = s.materials[s.inst_base:s.inst_base + s.inst_count * 28].reshape(-1, 28)
        ranges = sorted({(int(r[12]), int(r[13])) for r in inst})
        f.refit(ranges, refit_tlas=True)
        s2 = scene_with(s, vertex_data=vd)
        nodes = s.bvh_nodes.copy()
        for root, end in ranges:
            assert oracle.refit_blas(nodes, s2, root, end)
        assert np.array_equal(f.read(N.BUF_BVH_NODES, np.uint32), nodes)
        assert np.array_equal(f.read(N.BUF_VERTEX_DATA, np.float32), vd)
        wd = np.stack([m.T.reshape(-1) for m in worlds])
        mats = f.read(N.BUF_MATERIALS, np.float32)
        b64, pad = oracle.inst_bounds(mats, s.inst_base, s.inst_count, nodes, wd)
        assert np.array_equal(f.read(N.BUF_INST_BOUNDS, np.float32).reshape(-1, 6), b64)
        if traversal == N.TRAV_MT_ORDERED:
            assert np.array_equal(f.read(N.BUF_TLAS_SAH, np.uint32).reshape(-1, 8), oracle.build_sah_tlas(pad))
        f.tick(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
        scene2, _ = oracle_scene_after_edits(f, s, vd, nodes, wd)
        check_hits(f, scene2, [(w.grid_min, w.grid_size, w.res)], w.rays, w.depth_oct, w.frame_jitter,
                   w.hysteresis, label=f"refit {IDS[traversal]}")
    finally:
        f.close()


# ---------------------------------------------------------------- device skinning
@pytest.fixture(scope="module")
def small_dynamic():
    """The dynamic workload at reduced size: 6 tubes of 16 x 16 segments and 12 rigid
    instances in a 12 x 8 x 10 m box, one 8^3 cascade over the box, 64 rays."""
    return lambda: sg.dynamic(res=(8, 8, 8), n_cascades=1, rays=64, n_tubes=6, tube_seg=(16, 16), n_rigid=12,
                              n_lights=4, depth_oct=6, box=(12.0, 8.0, 10.0))


@pytest.mark.parametrize("traversal", MODES, ids=IDS)
def test_skinning_frames_match_oracle(small_dynamic, traversal):
    """Three frames of (skin -> rigid random walk -> tick): skinned vertices, refit BLASes,
    packed triangles, instance bounds and TLAS bit-exact, hits against the oracle."""
    w = small_dynamic()
    s = w.scene
    f = make_field(s, w.cascades, w.rays, w.depth_oct, traversal)
    try:
        worlds = [np.array(m) for m in w.slot_worlds]
        f.set_instance_transforms(np.arange(s.inst_count), worlds)
        f.set_skins(w.skin_sets, w.n_bones)
        for frame in range(3):
            t = frame / w.fps + 0.37
            bones = w.bones(t)
            print(f"case: skin frame {frame} {IDS[traversal]}", flush=True)
            f.skin(bones, update_tlas=False)
            slots, ws = w.rigid_step(sigma=0.3, sigma_rot=math.radians(10.0))
            for sl, m in zip(slots, ws):
                worlds[sl] = m
            f.set_instance_transforms(slots, ws)
            f.tick(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
        vd = s.vertex_data.copy().reshape(-1, 8)
        for vs, vc, br, be, j, wt in w.skin_sets:
            vd[vs:vs + vc, :6] = oracle.skin(s.vertex_data.reshape(-1, 8)[vs:vs + vc, :6], j, wt, bones)
        g_vd = f.read(N.BUF_VERTEX_DATA, np.float32).reshape(-1, 8)
        assert np.array_equal(g_vd[:, :3], vd[:, :3])
        assert np.array_equal(g_vd[:, 3:6], vd[:, 3:6])
        assert np.array_equal(g_vd[:, 6:], vd[:, 6:])
        s2 = scene_with(s, vertex_data=np.ascontiguousarray(vd.reshape(-1)))
        nodes = s.bvh_nodes.copy()
        for vs, vc, br, be, j, wt in w.skin_sets:
            assert oracle.refit_blas(nodes, s2, br, be)
        assert np.array_equal(f.read(N.BUF_BVH_NODES, np.uint32), nodes)
        tri = s.tri_index.reshape(-1, 3)
        pk = f.read(N.BUF_TRI_PACKED, np.float32).reshape(-1, 3, 4)
        p = vd[:, :3]
        assert np.array_equal(pk[:, 0, :3], p[tri[:, 0]])
        assert np.array_equal(pk[:, 1, :3], p[tri[:, 1]] - p[tri[:, 0]])
        assert np.array_equal(pk[:, 2, :3], p[tri[:, 2]] - p[tri[:, 0]])
        wd = np.stack([m.T.reshape(-1) for m in worlds])
        mats = f.read(N.BUF_MATERIALS, np.float32)
        b64, pad = oracle.inst_bounds(mats, s.inst_base, s.inst_count, nodes, wd)
        assert np.array_equal(f.read(N.BUF_INST_BOUNDS, np.float32).reshape(-1, 6), b64)
        if traversal == N.TRAV_MT_ORDERED:
            assert np.array_equal(f.read(N.BUF_TLAS_SAH, np.uint32).reshape(-1, 8), oracle.build_sah_tlas(pad))
        scene2, _ = oracle_scene_after_edits(f, s, vd, nodes, wd)
        check_hits(f, scene2, w.cascades, w.rays, w.depth_oct, w.frame_jitter, w.hysteresis,
                   label=f"skin {IDS[traversal]}")
        sk, rf, tl = f.edit_timing()
        print(f"edit timing: skin {sk:.4f} ms, BLAS refit {rf:.4f} ms, TLAS {tl:.4f} ms")
        assert sk > 0 and rf > 0 and tl > 0
    finally:
        f.close()


@pytest.mark.parametrize("traversal", MODES, ids=IDS)
def test_skin_after_rebuild_matches_oracle(small_dynamic, traversal):
    """Skin, rebuild the skinned BLASes on the device, skin again: the second refit runs on
    the plan the rebuild leaves behind (its depth table, deepest depth first), and must equal
    the oracle's refit of the rebuilt tree over the new pose bit for bit; the packed
    triangles follow the rebuilt triangle order."""
    w = small_dynamic()
    s = w.scene
    f = make_field(s, w.cascades, w.rays, w.depth_oct, traversal)
    try:
        f.set_instance_transforms(np.arange(s.inst_count), [np.array(m) for m in w.slot_worlds])
        f.set_skins(w.skin_sets, w.n_bones)
        f.skin(w.bones(0.61), update_tlas=False)
        ranges = [(int(br), int(be)) for _, _, br, be, _, _ in w.skin_sets]
        print(f"case: skin after rebuild {IDS[traversal]}, {len(ranges)} BLASes", flush=True)
        f.rebuild_blas(ranges)
        n1 = f.read(N.BUF_BVH_NODES, np.uint32)
        t1 = f.read(N.BUF_TRI_INDEX, np.uint32)
        m1 = f.read(N.BUF_TRI_MATERIAL, np.uint32)
        bones = w.bones(1.13)
        f.skin(bones, update_tlas=True)
        vd = s.vertex_data.copy().reshape(-1, 8)
        for vs, vc, br, be, j, wt in w.skin_sets:
            vd[vs:vs + vc, :6] = oracle.skin(s.vertex_data.reshape(-1, 8)[vs:vs + vc, :6], j, wt, bones)
        s2 = scene_with(s, vertex_data=np.ascontiguousarray(vd.reshape(-1)), tri_index=t1, tri_material=m1)
        nodes = n1.copy()
        for root, end in ranges:
            assert oracle.refit_blas(nodes, s2, root, end)
        assert np.array_equal(f.read(N.BUF_BVH_NODES, np.uint32), nodes)
        pk = f.read(N.BUF_TRI_PACKED, np.float32).reshape(-1, 3, 4)
        tri = t1.reshape(-1, 3)
        p = vd[:, :3]
        assert np.array_equal(pk[:, 0, :3], p[tri[:, 0]])
        assert np.array_equal(pk[:, 1, :3], p[tri[:, 1]] - p[tri[:, 0]])
    finally:
        f.close()


def test_set_skins_rejects_bad_sets(small_dynamic):
    w = small_dynamic()
    f = make_field(w.scene, w.cascades, w.rays, w.depth_oct, N.TRAV_MT_ORDERED)
    try:
        vs, vc, br, be, j, wt = w.skin_sets[0]
        bad_joint = j.copy()
        bad_joint[0, 0] = w.n_bones
        cases = [
            [w.skin_sets[0], w.skin_sets[0]],                       # overlapping sets
            [(vs, vc, br, be, bad_joint, wt)],                      # joint outside the palette
            [(vs, vc, br, be + 10 ** 7, j, wt)],                    # BLAS range out of bounds
            [(vs, w.scene.vertex_data.size, br, be, j, wt)],        # vertex range out of bounds
        ]
        for sets in cases:
            with pytest.raises(RuntimeError):
                f.set_skins(sets, w.n_bones)
        with pytest.raises(RuntimeError):
            f.skin(w.bones(0.0))                                    # nothing registered
        f.set_skins(w.skin_sets, w.n_bones)
        with pytest.raises(RuntimeError):
            f.skin(w.bones(0.0)[:-1])                               # palette size differs
    finally:
        f.close()


@pytest.mark.parametrize("n,kind", [(1, "random"), (2, "random"), (3, "random"), (37, "coincident"),
                                    (64, "inactive"), (2048, "random"), (3000, "random")])
def test_device_sah_tlas_edge_cases(n, kind):
    """The one-workgroup SAH build against the oracle's restatement, word for word, for a
    single instance, tiny sets, every instance at one point (no axis has centroid extent:
    middle splits), a quarter of the slots inactive (point boxes at their translation) and
    a few thousand instances (many levels of many nodes); 2048 slots is the largest build
    whose boxes and ids sit in LDS, 3000 reads them from global memory."""
    rng = np.random.default_rng(100 + n)
    b = sg.SceneBuilder()
    v, t = sg.box_mesh((0.0, 0.0, 0.0), (0.5, 0.3, 0.2), 0.0)
    mesh = b.add_mesh(v, t)
    worlds, active = [], []
    for i in range(n):
        if kind == "coincident":
            m = np.eye(4)
            m[:3, 3] = (1.0, 2.0, 3.0)
        else:
            m = sg.trs(rng.uniform(-50, 50, 3), sg.random_rotation(sg.SplitMix64(i + 7)), rng.uniform(0.2, 3.0))
        worlds.append(m)
        active.append(not (kind == "inactive" and i % 4 == 0))
        b.add_instance(mesh, m, active=active[-1])
    b.add_light(sg.point_light((0.0, 10.0, 0.0), intensity=10.0))
    scene = b.finalize(workload="tlas_edge")
    f = ProbeField((2, 2, 2), (-1.0, -1.0, -1.0), (2.0, 2.0, 2.0), rays=32, traversal=N.TRAV_MT_ORDERED)
    try:
        f.set_scene(scene)
        order = scene.tlas_order  # slot -> source instance
        print(f"case: device TLAS {kind} n={n}", flush=True)
        f.set_instance_transforms(np.arange(n), [worlds[order[s]] for s in range(n)],
                                  active=[active[order[s]] for s in range(n)])
        wd = np.stack([worlds[order[s]].T.reshape(-1) for s in range(n)])
        b64, pad = oracle.inst_bounds(f.read(N.BUF_MATERIALS, np.float32), scene.inst_base, n,
                                      f.read(N.BUF_BVH_NODES, np.uint32), wd)
        assert np.array_equal(f.read(N.BUF_INST_BOUNDS, np.float32).reshape(-1, 6), b64)
        sah = f.read(N.BUF_TLAS_SAH, np.uint32).reshape(-1, 8)
        assert np.array_equal(sah, oracle.build_sah_tlas(pad))
        oct_ = f.read(N.BUF_TLAS_OCT, np.uint32) if len(sah) <= 1023 else np.zeros(0, np.uint32)
        if len(sah) <= 1023:
            check_tlas_octants(sah, oct_.reshape(8, -1, 8))
        else:
            assert oct_.size == 0  # too large for copies: the closest-hit walk reads the SAH TLAS
    finally:
        f.close()


def tlas_octant_copies(sah):
    """The near-first copies of a walk-format TLAS: copy o orders each interior node's
    children by the projection of their box centres (min + max, in f32) on the octant
    direction (-1 on the axes whose bit is set), nearer first; positions follow the source's
    pre-order, miss = position + subtree size."""
    m = len(sah)
    f = sah.view(np.float32)
    out = np.zeros((8, m, 8), np.uint32)
    for o in range(8):
        sgn = np.float32([-1.0 if o & (1 << a) else 1.0 for a in range(3)])
        q = np.zeros(m, np.int64)
        for i in range(m):
            if sah[i, 7] != 0xFFFFFFFF:
                continue
            l, r = i + 1, int(sah[i + 1, 3])
            cl = (f[l, 0] + f[l, 4]) * sgn[0] + (f[l, 1] + f[l, 5]) * sgn[1]
            cl = np.float32(cl) + np.float32((f[l, 2] + f[l, 6]) * sgn[2])
            cr = (f[r, 0] + f[r, 4]) * sgn[0] + (f[r, 1] + f[r, 5]) * sgn[1]
            cr = np.float32(cr) + np.float32((f[r, 2] + f[r, 6]) * sgn[2])
            size_l, size_r = r - l, int(sah[i, 3]) - r
            near_l = cl <= cr
            q[l] = q[i] + 1 if near_l else q[i] + 1 + size_r
            q[r] = q[i] + 1 + size_l if near_l else q[i] + 1
        for i in range(m):
            row = sah[i].copy()
            row[3] = q[i] + (int(sah[i, 3]) - i)
            out[o, q[i]] = row
    return out


def check_tlas_octants(sah, got):
    want = tlas_octant_copies(sah)
    assert np.array_equal(got, want)
    # each copy is a valid threaded layout over the same leaves
    for o in range(8):
        c = got[o]
        m = len(c)
        assert sorted(c[c[:, 7] != 0xFFFFFFFF, 7].tolist()) == sorted(sah[sah[:, 7] != 0xFFFFFFFF, 7].tolist())
        assert np.all(c[:, 3] > np.arange(m)) and np.all(c[:, 3] <= m)


# ---------------------------------------------------------------- device BLAS rebuild
@pytest.mark.parametrize("traversal", MODES, ids=IDS)
def test_rebuild_blas_matches_oracle(instanced_small, traversal):
    """After a deformation that a refit would serve badly (axes rotated, scaled), every
    instanced BLAS is rebuilt on the device: same node count; each range's triangle rows are
    a permutation of its old rows; the oracle's refit of the rebuilt tree reproduces the
    device's bounds bit for bit (so the threaded layout is valid and the bounds tight);
    hits on the rebuilt scene match the oracle's."""
    w = instanced_small
    s = w.scene
    f = make_field(s, [(w.grid_min, w.grid_size, w.res)], w.rays, w.depth_oct, traversal)
    try:
        worlds = slot_worlds_from_rows(s)
        f.set_instance_transforms(np.arange(s.inst_count), worlds)
        vd = s.vertex_data.copy().reshape(-1, 8)
        vd[:, :3] = (vd[:, [1, 2, 0]] * np.float32([1.6, 0.7, 1.2])).astype(np.float32)
        vd = np.ascontiguousarray(vd.reshape(-1))
        f.update_ranges(N.BUF_VERTEX_DATA, [0, vd.size], vd)
        inst = s.materials[s.inst_base:s.inst_base + s.inst_count * 28].reshape(-1, 28)
        ranges = sorted({(int(r[12]), int(r[13])) for r in inst})
        print(f"case: rebuild {IDS[traversal]}, {len(ranges)} BLASes", flush=True)
        f.rebuild_blas(ranges)
        nodes = f.read(N.BUF_BVH_NODES, np.uint32)
        ti = f.read(N.BUF_TRI_INDEX, np.uint32)
        tm = f.read(N.BUF_TRI_MATERIAL, np.uint32)
        assert nodes.size == s.bvh_nodes.size
        old_rows = np.concatenate([s.tri_index.reshape(-1, 3), s.tri_material[:, None]], 1)
        new_rows = np.concatenate([ti.reshape(-1, 3), tm[:, None]], 1)
        s2 = scene_with(s, vertex_data=vd, tri_index=ti, tri_material=tm)
        nu = nodes.reshape(-1, 8)
        for root, end in ranges:
            pay = nu[root:end, 7]
            leaf = pay[pay != 0xFFFFFFFF]
            t0, t1 = int((leaf & 0xFFFFFF).min()), int(((leaf & 0xFFFFFF) + (leaf >> 24)).max())
            cover = np.zeros(len(old_rows), np.int32)
            for p in leaf:
                cover[(p & 0xFFFFFF):(p & 0xFFFFFF) + (p >> 24)] += 1
            assert np.all(cover[t0:t1] == 1)
            a = old_rows[t0:t1]
            b_ = new_rows[t0:t1]
            assert np.array_equal(a[np.lexsort(a.T[::-1])], b_[np.lexsort(b_.T[::-1])])
            refit = nodes.copy()
            assert oracle.refit_blas(refit, s2, root, end)
            assert np.array_equal(refit.reshape(-1, 8)[root:end], nu[root:end])
        wd = np.stack([m.T.reshape(-1) for m in worlds])
        mats = f.read(N.BUF_MATERIALS, np.float32)
        b64, pad = oracle.inst_bounds(mats, s.inst_base, s.inst_count, nodes, wd)
        assert np.array_equal(f.read(N.BUF_INST_BOUNDS, np.float32).reshape(-1, 6), b64)
        if traversal == N.TRAV_MT_ORDERED:
            assert np.array_equal(f.read(N.BUF_TLAS_SAH, np.uint32).reshape(-1, 8), oracle.build_sah_tlas(pad))
        f.tick(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
        scene2, _ = oracle_scene_after_edits(f, s2, vd, nodes, wd)
        check_hits(f, scene2, [(w.grid_min, w.grid_size, w.res)], w.rays, w.depth_oct, w.frame_jitter,
                   w.hysteresis, label=f"rebuild {IDS[traversal]}")
        # a later refit of the rebuilt ranges (cached plans were invalidated) keeps the bounds
        f.refit(ranges, refit_tlas=True)
        assert np.array_equal(f.read(N.BUF_BVH_NODES, np.uint32), nodes)
        sk, rb, tl = f.edit_timing()
        print(f"rebuild timing: BLAS refit {rb:.4f} ms, TLAS {tl:.4f} ms")
        # the scratch is kept between rebuilds: a smaller rebuild, then the full one again,
        # reproduce the tree (the triangles are already in Morton order and the sort is
        # stable, so a rebuild of unchanged vertices is idempotent)
        f.rebuild_blas(ranges[:1])
        assert np.array_equal(f.read(N.BUF_BVH_NODES, np.uint32), nodes)
        assert np.array_equal(f.read(N.BUF_TRI_INDEX, np.uint32), ti)
        t0 = time.perf_counter()
        f.rebuild_blas(ranges)
        wall = (time.perf_counter() - t0) * 1e3
        assert np.array_equal(f.read(N.BUF_BVH_NODES, np.uint32), nodes)
        assert np.array_equal(f.read(N.BUF_TRI_INDEX, np.uint32), ti)
        assert np.array_equal(f.read(N.BUF_TRI_MATERIAL, np.uint32), tm)
        print(f"rebuild of {len(ranges)} BLASes with kept scratch: {wall:.3f} ms host wall")
    finally:
        f.close()
