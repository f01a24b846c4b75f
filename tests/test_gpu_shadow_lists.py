"""The shadow walk's light-space candidate lists (run on an MI355X with -m gpu).

The lists replace the TLAS walk of NEE shadow rays (traverseAny, js/spectral_traverse.js:
480-569): per light, a cube map of cells around a point or spot light (a grid across the
direction of a directional light) lists the instances whose padded box reaches the cell,
and a shadow ray walks its cell's list. Two properties are checked:

- conservative: for segments from random origins toward every light (within a ranged
  light's range), every instance whose world box the segment meets in FP64 is in the list
  of the segment's cell, computed here by a float32 restatement of the walk's cell lookup;
  the lists are also selective (far shorter than the instance count), and each holds its
  boxes nearest the light first (the order cannot change a mask; it makes a blocked ray
  stop sooner);
- same answers: the per-ray light visibility masks of a field that walks the lists equal,
  bit for bit, those of the same field walking the TLAS, on the instanced workload (also
  after instance edits, which rebuild the lists on the device), the room workload
  (directional light, one instance) and seeded random scenes with mirrored and inactive
  instances and point, ranged, spot and directional lights.
"""
from __future__ import annotations

import numpy as np
import pytest

from speedball_amd import _native as N
from speedball_amd import scenegen as sg
from speedball_amd.field import ProbeField
from tests.test_gpu_random_scenes import random_workload

pytestmark = pytest.mark.gpu

FACE = 16        # cube-map cells per face edge
DIR_GRID = 32    # directional grid cells per axis
CELLS = 6 * FACE * FACE
RECS = CELLS + 2


def host_cell(pj, o):
    """Cell record index of shadow rays from origins o (float32 [n, 3]) for a light whose
    projection record is pj (float32 [4, 4]), restating the walk's lookup in float32."""
    f32 = np.float32
    kind = pj[0, 0]
    if kind == 0.0:
        v = (o - pj[1, :3]).astype(f32)
        a = np.abs(v)
        fx = (a[:, 0] >= a[:, 1]) & (a[:, 0] >= a[:, 2])
        fy = ~fx & (a[:, 1] >= a[:, 2])
        m = np.where(fx, a[:, 0], np.where(fy, a[:, 1], a[:, 2]))
        major = np.where(fx, v[:, 0], np.where(fy, v[:, 1], v[:, 2]))
        aa = np.where(fx, v[:, 1], v[:, 0])
        bb = np.where(fx | fy, v[:, 2], v[:, 1])
        face = np.where(fx, 0, np.where(fy, 2, 4)) + (major < 0)
        h = f32(FACE / 2)
        with np.errstate(divide="ignore", invalid="ignore"):
            sc = h / m
            iu = np.nan_to_num(np.clip(aa * sc + h, 0, FACE - 1), nan=0.0).astype(np.int64)
            iv = np.nan_to_num(np.clip(bb * sc + h, 0, FACE - 1), nan=0.0).astype(np.int64)
        return (face * FACE + iv) * FACE + iu
    if kind == 1.0:
        fu = ((o @ pj[1, :3]).astype(f32) - pj[1, 3]) * pj[3, 0]
        fv = ((o @ pj[2, :3]).astype(f32) - pj[2, 3]) * pj[3, 1]
        inside = (fu >= 0) & (fu < DIR_GRID) & (fv >= 0) & (fv < DIR_GRID)
        c = np.clip(fv, 0, DIR_GRID - 1).astype(np.int64) * DIR_GRID + np.clip(fu, 0, DIR_GRID - 1).astype(np.int64)
        return np.where(inside, c, CELLS + 1)
    return np.full(len(o), CELLS, np.int64)


def segment_meets_boxes(p, q, boxes):
    """[n_boxes] bool: the segment p -> q (float64) meets each box (slab test, closed)."""
    d = q - p
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        t0 = (boxes[:, :3] - p) * inv
        t1 = (boxes[:, 3:] - p) * inv
    lo = np.where(np.isnan(t0), -np.inf, np.minimum(t0, t1))
    hi = np.where(np.isnan(t0), np.inf, np.maximum(t0, t1))
    # an axis along which the segment does not move: inside the slab or not at all
    flat = d == 0
    inside = (p >= boxes[:, :3]) & (p <= boxes[:, 3:])
    lo = np.where(flat, np.where(inside, -np.inf, np.inf), lo)
    hi = np.where(flat, np.where(inside, np.inf, -np.inf), hi)
    tn = np.maximum(lo.max(1), 0.0)
    tf = np.minimum(hi.min(1), 1.0)
    return tn <= tf


def check_lists(f, scene, rng, n_origins=3000):
    lights = scene.lights.view(np.float32).reshape(-1, 17)
    n_inst = scene.inst_count
    proj = f.read(N.BUF_SHADOW_PROJ, np.float32).reshape(-1, 4, 4)
    cells = f.read(N.BUF_SHADOW_CELLS, np.uint32).reshape(-1, RECS, 2)
    lst = f.read(N.BUF_SHADOW_LIST, np.uint16)
    boxes = f.read(N.BUF_INST_BOUNDS, np.float32).reshape(-1, 6).astype(np.float64)
    assert proj.shape[0] == len(lights) and boxes.shape[0] == n_inst
    lo, hi = boxes[:, :3].min(0), boxes[:, 3:].max(0)
    span = hi - lo
    checked, lengths = 0, []
    for li, L in enumerate(lights):
        kind = proj[li, 0, 0]
        if L[0] > 2.5:
            assert kind == 2.0
            continue
        # origins: uniform over the scene, and on the faces of random instance boxes
        u = rng.uniform(-0.05, 1.05, (n_origins, 3)) * span + lo
        bi = rng.integers(0, n_inst, n_origins)
        on = boxes[bi, :3] + rng.uniform(0, 1, (n_origins, 3)) * (boxes[bi, 3:] - boxes[bi, :3])
        ax = rng.integers(0, 3, n_origins)
        side = rng.integers(0, 2, n_origins)
        on[np.arange(n_origins), ax] = np.where(side == 1, boxes[bi, 3 + ax], boxes[bi, ax])
        origins = np.concatenate([u, on]).astype(np.float32)
        if kind == 0.0:
            pos = L[1:4].astype(np.float64)
            ends = np.repeat(pos[None], len(origins), 0)
            if L[10] > 0:  # ranged: only origins within range queue shadow rays
                keep = np.linalg.norm(origins - pos, axis=1) < L[10]
                origins, ends = origins[keep], ends[keep]
        else:
            w = -L[4:7].astype(np.float64)
            w /= np.linalg.norm(w)
            ends = origins.astype(np.float64) + 1e4 * w
        c = host_cell(proj[li], origins)
        for j in range(len(origins)):
            first, end = cells[li, c[j]]
            listed = set(lst[first:end].tolist())
            lengths.append(end - first)
            need = np.nonzero(segment_meets_boxes(origins[j].astype(np.float64), ends[j], boxes))[0]
            missing = [int(i) for i in need if int(i) not in listed]
            assert not missing, (f"light {li} (type {L[0]:.0f}) origin {origins[j]} cell {c[j]}: "
                                 f"instances {missing} met but not listed")
            checked += 1
    return checked, float(np.mean(lengths)) if lengths else 0.0


def padded_boxes(bounds):
    """The padded world boxes the lists are built from (float32 [n, 6]): each axis widened by
    1e-5 of (extent + |lo| + |hi|) plus 1e-6, as k_inst_bounds pads them."""
    f32 = np.float32
    lo, hi = bounds[:, :3].astype(f32), bounds[:, 3:].astype(f32)
    pad = f32(1e-5) * (np.abs(hi - lo) + np.abs(lo) + np.abs(hi)) + f32(1e-6)
    return np.concatenate([lo - pad, hi + pad], 1)


def order_keys(pj, boxes):
    """Sort key of every box for a light's projection record pj (float64): the squared
    distance from a point or spot light to the box, or minus the box centre's height along
    the direction toward a directional light (the cross product of the grid's axes)."""
    b = boxes.astype(np.float64)
    if pj[0, 0] == 0.0:
        lp = pj[1, :3].astype(np.float64)
        d = np.maximum(np.maximum(b[:, :3] - lp, lp - b[:, 3:]), 0.0)
        return (d * d).sum(1)
    if pj[0, 0] == 1.0:
        w = np.cross(pj[1, :3].astype(np.float64), pj[2, :3].astype(np.float64))
        return -(0.5 * (b[:, :3] + b[:, 3:])) @ w
    return np.zeros(len(b))


def check_order(f, scene):
    """Every cell list holds its slots nearest the light first (ties by slot). Returns the
    adjacent pairs checked, how many of them step to a clearly larger key, and how many
    lists are not in slot order."""
    n_inst = scene.inst_count
    proj = f.read(N.BUF_SHADOW_PROJ, np.float32).reshape(-1, 4, 4)
    cells = f.read(N.BUF_SHADOW_CELLS, np.uint32).reshape(-1, RECS, 2)
    lst = f.read(N.BUF_SHADOW_LIST, np.uint16)
    boxes = padded_boxes(f.read(N.BUF_INST_BOUNDS, np.float32).reshape(-1, 6))
    pairs = rising = unsorted_by_slot = 0
    for li in range(proj.shape[0]):
        if proj[li, 0, 0] > 1.5:
            continue
        key = order_keys(proj[li], boxes)
        for c in range(CELLS):
            first, end = (int(x) for x in cells[li, c])
            if end - first < 2 or first < n_inst + 1:  # one entry, or the identity list
                continue
            e = lst[first:end].astype(np.int64)
            ka, kb = key[e[:-1]], key[e[1:]]
            tol = 1e-4 * (np.abs(ka) + np.abs(kb)) + 1e-5
            bad = (ka > kb + tol) | ((ka == 0.0) & (kb == 0.0) & (e[:-1] >= e[1:]))
            assert not bad.any(), (f"light {li} cell {c}: slots {e[:-1][bad]} (keys {ka[bad]}) listed before "
                                   f"slots {e[1:][bad]} (keys {kb[bad]})")
            pairs += len(e) - 1
            rising += int((kb > ka + tol).sum())
            unsorted_by_slot += int((np.diff(e) < 0).any())
    return pairs, rising, unsorted_by_slot


def test_lists_walk_the_nearest_boxes_first():
    """A blocked shadow ray stops at its first occluder, so each list holds the boxes nearest
    its light first: checked on every cell list of the instanced workload (point lights),
    seeded random scenes (point, ranged, spot and directional lights) and 900 instances."""
    cases = [("instanced", sg.instanced(res=(8, 4, 6), rays=64))]
    cases += [(f"random{s}", random_workload(s)) for s in (11, 37, 73)]
    cases += [("many900", many_instances())]
    for name, w in cases:
        print(f"case: {name}, {w.scene.inst_count} instances", flush=True)
        f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=w.depth_oct)
        try:
            f.set_scene(w.scene)
            pairs, rising, unsorted = check_order(f, w.scene)
        finally:
            f.close()
        print(f"{name}: {pairs} adjacent pairs in order ({rising} rising), {unsorted} lists not in slot order")
        if w.scene.inst_count > 8:
            assert rising > 0 and unsorted > 0


def test_lists_hold_every_reachable_instance():
    rng = np.random.default_rng(7)
    w = sg.instanced(res=(8, 4, 6), rays=64)
    print(f"case: instanced scene, {w.scene.inst_count} instances, {len(w.scene.lights) // 17} lights", flush=True)
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=w.depth_oct, traversal=N.TRAV_MT_ORDERED)
    try:
        f.set_scene(w.scene)
        n, mean_len = check_lists(f, w.scene, rng, 1500)
    finally:
        f.close()
    print(f"instanced: {n} segments checked, mean list length {mean_len:.2f} of {w.scene.inst_count} instances")
    assert n > 10000
    assert mean_len < w.scene.inst_count / 8  # selective, not the every-instance fallback
    for seed in (11, 59, 73):
        w = random_workload(seed)
        types = sorted(set(int(t) for t in w.scene.lights.view(np.float32).reshape(-1, 17)[:, 0]))
        print(f"case: random scene seed {seed}, light types {types}", flush=True)
        f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=w.depth_oct)
        try:
            f.set_scene(w.scene)
            n, mean_len = check_lists(f, w.scene, rng, 800)
        finally:
            f.close()
        print(f"random {seed}: {n} segments checked, mean list length {mean_len:.2f} of {w.scene.inst_count}")
        assert n > 1000


def _masks(w, lists, traversal, edit=None):
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=w.depth_oct, traversal=traversal,
                   shadow_lists=lists)
    try:
        f.set_scene(w.scene)
        assert (f.nbytes(N.BUF_SHADOW_LIST) > 0) == lists
        out = []
        f.tick(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
        out.append(f.read(N.BUF_SHADOW_VIS, np.uint64))
        if edit is not None:
            edit(f)
            f.tick(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
            out.append(f.read(N.BUF_SHADOW_VIS, np.uint64))
        return out
    finally:
        f.close()


def _move_instances(w):
    rng = np.random.default_rng(3)
    mats = w.scene.materials.view(np.float32)
    slots = rng.choice(w.scene.inst_count, 32, replace=False)
    worlds = []
    for s in slots:
        rec = mats[w.scene.inst_base + 28 * s: w.scene.inst_base + 28 * s + 12].reshape(3, 4).astype(np.float64)
        M = np.linalg.inv(np.vstack([rec, [0, 0, 0, 1]]))
        M[:3, 3] += rng.uniform(-3, 3, 3)
        worlds.append(M)
    return lambda f: f.set_instance_transforms(slots, worlds)


@pytest.mark.parametrize("case", ["instanced", "room", "random11", "random37", "random73"])
def test_lists_give_the_tlas_walks_masks(case):
    if case == "instanced":
        w = sg.instanced(res=(16, 8, 12), rays=128)
        edit = _move_instances(w)
    elif case == "room":
        w, edit = sg.room_boxes(res=(12, 12, 12), rays=64), None
    else:
        w, edit = random_workload(int(case[6:])), None
    for trav in (N.TRAV_MT_ORDERED, N.TRAV_MT_COMPAT):
        print(f"case: {case}, traversal {trav}", flush=True)
        a = _masks(w, True, trav, edit)
        b = _masks(w, False, trav, edit)
        for i, (x, y) in enumerate(zip(a, b)):
            bits = int(np.unpackbits((x ^ y).view(np.uint8)).sum())
            ones = int(np.unpackbits(y.view(np.uint8)).sum())
            print(f"{case} traversal {trav} update {i}: {ones} unblocked shadow rays, {bits} bits differ")
            assert ones > 0
            assert bits == 0


def test_walk_tuning_switches_the_lists():
    """The lists are a walk-tuning choice (default on for a cache-resident scene): switching
    them off, editing instances while they are off (the builds are skipped) and on again
    (the stale lists are rebuilt) keeps the masks equal to the TLAS walk's."""
    w = sg.instanced(res=(12, 6, 8), rays=64)
    edit = _move_instances(w)
    kw = dict(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=w.depth_oct, traversal=N.TRAV_MT_ORDERED)
    try:
        f.set_scene(w.scene)
        assert f.walk_tuning()["shadow_lists"] == 1
        f.tick(**kw)
        on = f.read(N.BUF_SHADOW_VIS, np.uint64)
        f.set_walk_tuning(shadow_lists=0)
        f.tick(**kw)
        off = f.read(N.BUF_SHADOW_VIS, np.uint64)
        edit(f)  # lists not rebuilt while off
        f.tick(**kw)
        off_edited = f.read(N.BUF_SHADOW_VIS, np.uint64)
        f.set_walk_tuning(shadow_lists=1)  # stale lists rebuilt
        f.tick(**kw)
        on_edited = f.read(N.BUF_SHADOW_VIS, np.uint64)
    finally:
        f.close()
    for name, a, b in (("before the edit", on, off), ("after the edit", on_edited, off_edited)):
        bits = int(np.unpackbits((a ^ b).view(np.uint8)).sum())
        print(f"{name}: {int(np.unpackbits(b.view(np.uint8)).sum())} unblocked shadow rays, {bits} bits differ")
        assert bits == 0
    assert int(np.unpackbits((on ^ on_edited).view(np.uint8)).sum()) > 0  # the edit moved shadows


def many_instances(n=900, seed=5):
    """n instances of three small meshes (a box, a sphere, a torus) over a 30 x 10 x 20 m
    box: more slots than the walk stages in LDS, so the list walk reads its boxes and
    instance records from global memory. Point, ranged, spot and directional lights."""
    from speedball_amd.scene import SceneBuilder, directional_light, point_light, spot_light, uber_material
    rng = np.random.default_rng(seed)
    box = np.array([30.0, 10.0, 20.0])
    b = SceneBuilder()
    meshes = []
    for kind, (v, t) in enumerate((sg.box_mesh((0, 0, 0), (0.4, 0.3, 0.5), 0.3), sg.uv_sphere(6, 12, 0.5),
                                   sg.torus(16, 8, 0.5, 0.15))):
        mat = b.add_material(uber_material((0.6, 0.5, 0.4), side=2 if kind == 2 else 0))
        meshes.append(b.add_mesh(np.asarray(v, np.float32), np.asarray(t, np.uint32),
                                 tri_material=np.full(len(t), mat, np.uint32)))
    for i in range(n):
        M = np.eye(4)
        a = rng.uniform(0, 2 * np.pi)
        M[:3, :3] = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]]) * rng.uniform(0.5, 1.5)
        M[:3, 3] = rng.uniform(0, 1, 3) * box
        b.add_instance(meshes[i % 3], M, active=bool(rng.uniform() > 0.05))
    b.add_light(point_light((15.0, 9.0, 10.0), intensity=40.0))
    b.add_light(point_light((5.0, 5.0, 5.0), intensity=20.0, distance=6.0, decay=1.5))
    b.add_light(spot_light((25.0, 8.0, 15.0), (20.0, 0.0, 10.0), intensity=60.0, angle=0.7, penumbra=0.3))
    b.add_light(directional_light((15.0, 30.0, 10.0), (14.0, 0.0, 12.0), intensity=1.0))
    scene = b.finalize(workload="many_instances")
    return sg.Workload("many_instances", scene, (0.0, 0.0, 0.0), tuple(box.tolist()), (10, 4, 7), 64, 6, 0.37)


def test_lists_from_global_memory_give_the_tlas_walks_masks():
    w = many_instances()
    print(f"case: {w.scene.inst_count} instances (list boxes and records from global memory)", flush=True)
    assert w.scene.inst_count * 96 > 64 * 1024
    for trav in (N.TRAV_MT_ORDERED, N.TRAV_MT_COMPAT):
        a = _masks(w, True, trav)[0]
        b = _masks(w, False, trav)[0]
        bits = int(np.unpackbits((a ^ b).view(np.uint8)).sum())
        ones = int(np.unpackbits(b.view(np.uint8)).sum())
        print(f"traversal {trav}: {ones} unblocked shadow rays, {bits} bits differ")
        assert ones > 0 and bits == 0
    rng = np.random.default_rng(11)
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=w.depth_oct)
    try:
        f.set_scene(w.scene)
        n, mean_len = check_lists(f, w.scene, rng, 400)
    finally:
        f.close()
    print(f"{n} segments checked, mean list length {mean_len:.2f} of {w.scene.inst_count}")
    assert n > 1000
