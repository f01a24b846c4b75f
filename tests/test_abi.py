"""C ABI and scene-builder checks that need no GPU.

* libsbgi.so loads and exports every function include/sbgi.h declares;
* the builder emits the reference's threaded layouts (left child at i+1, miss = end of
  subtree, every triangle in exactly one leaf, bounds enclose children);
* the TLAS equals a direct restatement of buildTlasRecords (js/spectral_scene.js:973-1003);
* transform updates equal a fresh build of the same instances;
* argument errors are reported with the reference's fail-closed codes.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

from speedball_amd import _native as N
from speedball_amd import scenegen as sg
from speedball_amd.scene import SceneBuilder, uber_material, point_light

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "sbgi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sbgi_\w+)\s*\(", src)))


def test_library_exports_every_header_symbol():
    L = N.lib()
    decl = header_functions()
    assert len(decl) >= 25
    for name in decl:
        assert hasattr(L, name), f"{name} missing from libsbgi.so"
    assert sorted(N.EXPORTS) == decl
    assert N.lib().sbgi_version().startswith(b"sbgi")


def test_default_params_match_reference_uniforms():
    p = N.default_tick_params()
    # js/gi_probes.js:1198-1255 defaults, gated jitter mode
    assert p.depth_sharpness == 7 and p.radiance_clamp == 8 and p.filter_strength == 1
    assert abs(p.filter_smooth - 0.5) < 1e-7 and abs(p.hysteresis - 0.6) < 1e-7
    assert abs(p.temp_change_h_drop - 0.3) < 1e-7 and p.temp_clamp_sigma == 6 and p.temp_change_sigma1 == 2.5
    assert p.dbg_cosine_power == 1 and abs(p.dbg_min_change_h - 0.55) < 1e-7


def test_create_without_device_fails_loudly():
    d = N.FieldDesc()
    d.rays_per_probe = 64
    d.cascade_count = 1
    d.depth_oct_res = 6
    d.cascades[0].res[:] = (2, 2, 2)
    h = C.c_void_p()
    rc = N.lib().sbgi_create(C.byref(d), C.byref(h))
    assert rc < 0 and not h.value
    d.rays_per_probe = 70  # not a multiple of 16
    assert N.lib().sbgi_create(C.byref(d), C.byref(h)) == N.SBGI_E_INVALID
    assert b"multiple of 16" in N.lib().sbgi_last_error(None)


def check_threaded(nodes_u32, root, end, tri_count=None):
    nu = nodes_u32.reshape(-1, 8)
    nf = nodes_u32.view(np.float32).reshape(-1, 8)
    seen = []

    def walk(i):
        # returns end of subtree
        if nu[i, 7] != 0xFFFFFFFF:
            off, cnt = nu[i, 7] & 0xFFFFFF, nu[i, 7] >> 24
            assert 1 <= cnt <= 8
            seen.extend(range(off, off + cnt))
            assert nu[i, 6] == i + 1
            return i + 1
        l = i + 1
        le = walk(l)
        r = nu[l, 6]
        assert r == le, "left child's miss must be the right child"
        re_ = walk(r)
        assert nu[i, 6] == re_
        for c in (l, r):
            assert np.all(nf[i, 0:3] <= nf[c, 0:3]) and np.all(nf[i, 3:6] >= nf[c, 3:6])
        return re_

    assert walk(root) == end
    return seen


def test_blas_layout_invariants(instanced_small):
    s = instanced_small.scene
    starts = sorted({int(s.materials[s.inst_base + i * 28 + 12]) for i in range(s.inst_count)})
    ends = sorted({int(s.materials[s.inst_base + i * 28 + 13]) for i in range(s.inst_count)})
    tris = []
    for r, e in zip(starts, ends):
        tris += check_threaded(s.bvh_nodes, r, e)
    assert sorted(tris) == list(range(s.tri_count)), "every pooled triangle in exactly one leaf"
    # leaf bounds enclose their triangles
    nu = s.bvh_nodes.reshape(-1, 8)
    nf = s.bvh_nodes.view(np.float32).reshape(-1, 8)
    v = s.vertex_data.reshape(-1, 8)[:, :3]
    ti = s.tri_index.reshape(-1, 3)
    for i in np.nonzero(nu[:, 7] != 0xFFFFFFFF)[0][:500]:
        off, cnt = nu[i, 7] & 0xFFFFFF, nu[i, 7] >> 24
        p = v[ti[off:off + cnt].reshape(-1)]
        assert np.array_equal(p.min(0), nf[i, 0:3]) and np.array_equal(p.max(0), nf[i, 3:6])


def reference_tlas(aabbs):
    """buildTlasRecords restated directly (stable sort, leaf <= 2, midpoint split)."""
    order = list(range(len(aabbs)))
    recs = []

    def emit(lo, hi):
        rec = {"bx": [np.inf] * 3 + [-np.inf] * 3, "leaf": False, "off": 0, "cnt": 0, "miss": 0}
        recs.append(rec)
        for i in range(lo, hi):
            a = aabbs[order[i]]
            for k in range(3):
                rec["bx"][k] = min(rec["bx"][k], a[0][k])
                rec["bx"][3 + k] = max(rec["bx"][3 + k], a[1][k])
        if hi - lo <= 2:
            rec.update(leaf=True, off=lo, cnt=hi - lo)
        else:
            ex, ey, ez = (rec["bx"][3] - rec["bx"][0], rec["bx"][4] - rec["bx"][1], rec["bx"][5] - rec["bx"][2])
            axis = 0 if (ex >= ey and ex >= ez) else (1 if ey >= ez else 2)
            order[lo:hi] = sorted(order[lo:hi], key=lambda i: (aabbs[i][0][axis] + aabbs[i][1][axis]) * 0.5)
            mid = (lo + hi) >> 1
            emit(lo, mid)
            emit(mid, hi)
        rec["miss"] = len(recs)

    emit(0, len(aabbs))
    return recs, order


def test_tlas_matches_reference_restatement():
    rng = np.random.default_rng(11)
    b = SceneBuilder()
    b.add_material(uber_material())
    v, t = sg.box_mesh((0, 0, 0), (0.5, 0.5, 0.5), 0.0)
    m = b.add_mesh(v, t)
    aabbs = []
    for i in range(37):
        pos = rng.uniform(-20, 20, 3)
        s = rng.uniform(0.5, 2.0)
        W = np.eye(4)
        W[:3, :3] *= s
        W[:3, 3] = pos
        b.add_instance(m, W)
        aabbs.append((pos - 0.5 * s, pos + 0.5 * s))
    b.add_light(point_light((0, 0, 0)))
    sc = b.finalize()
    recs, order = reference_tlas(aabbs)
    assert list(sc.tlas_order) == order
    tl = sc.materials[sc.tlas_base:].reshape(-1, 12)
    assert len(recs) == sc.tlas_node_count
    for i, r in enumerate(recs):
        assert tl[i, 6] == r["miss"]
        assert tl[i, 8] == (r["cnt"] if r["leaf"] else 0)
        assert np.allclose(tl[i, 0:6], np.float32(r["bx"]), rtol=1e-6)


def test_set_transform_equals_fresh_build():
    rng = np.random.default_rng(12)
    v, t = sg.uv_sphere(6, 12)

    def build(worlds):
        b = SceneBuilder()
        b.add_material(uber_material())
        m = b.add_mesh(v, t)
        for W in worlds:
            b.add_instance(m, W)
        b.add_light(point_light((0, 5, 0)))
        return b

    worlds = [sg.trs(rng.uniform(-10, 10, 3), sg.random_rotation(sg.SplitMix64(i)), 1.0) for i in range(9)]
    b = build(worlds)
    s0 = b.finalize()
    moved = worlds[4].copy()
    moved[:3, 3] += (0.3, -0.2, 0.1)  # small move keeps the frozen partition identical
    slot = b.set_transform(4, moved)
    assert s0.tlas_order[slot] == 4
    s1 = b.finalize()
    worlds2 = list(worlds)
    worlds2[4] = moved
    fresh = build(worlds2).finalize()
    if list(fresh.tlas_order) == list(s0.tlas_order):
        assert np.array_equal(s1.materials, fresh.materials)


def test_builder_rejects_bad_input():
    b = SceneBuilder()
    with pytest.raises(N.SbgiError):
        b.add_mesh(np.zeros((3, 3), np.float32), np.array([[0, 1, 5]], np.uint32))
    for _ in range(64):
        b.add_light(point_light((0, 0, 0)))
    with pytest.raises(N.SbgiError) as e:
        b.add_light(point_light((0, 0, 0)))
    assert e.value.code == N.SBGI_E_CAPACITY


def test_scenegen_workloads_sizes():
    w = sg.room_boxes()
    assert w.scene.tri_count == 780 and w.probes == 4096 and w.rays == 64
    assert w.scene.lights.size == 2 * 17
