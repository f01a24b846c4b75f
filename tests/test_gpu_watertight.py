"""SBGI_TRAV_WATERTIGHT on the GPU (run on an MI355X with -m gpu).

The watertight mode replaces Moller-Trumbore by the watertight triangle test (Woop, Benthin
and Wald 2013) in both walks. Against the oracle's FP32 restatement of the same test its
hit ids are identical except near ties (the walk order of the ordered layouts decides ties)
and its radiance is within the FP32 bars; against Moller-Trumbore (the reference's test) its
hits differ only at triangle edges and near ties, counted and reported; and on a closed mesh
no ray aimed exactly at a shared vertex or edge leaks through it, where Moller-Trumbore lets
some through (also counted)."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from speedball_amd import _native as N
from speedball_amd import scenegen as sg
from speedball_amd.field import ProbeField
from tests.parity import classify_mismatches, rel_err, summary
from tests.test_oracle_pins import _edge_and_vertex_rays, _mesh_scene

pytestmark = pytest.mark.gpu


def _run(w, traversal):
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=w.depth_oct, keep_hit_ids=True,
                   traversal=traversal)
    f.set_scene(w.scene)
    try:
        f.tick(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
        return f.read_rays(), f.read_hit_ids()
    finally:
        f.close()


@pytest.mark.parametrize("name", ["room", "instanced"])
def test_watertight_field_matches_its_restatement(room, instanced_small, name):
    w = room if name == "room" else instanced_small
    print(f"case: watertight field, {name}", flush=True)
    g_rays, g_hits = _run(w, N.TRAV_WATERTIGHT)
    m_rays, m_hits = _run(w, N.TRAV_MT_ORDERED)
    oracle.set_watertight(True, "32")
    try:
        o = oracle.OracleField(w.scene, w.grid_min, w.grid_size, w.res, w.rays, w.depth_oct, prec="32")
        o.update(o.params(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis))
        r_rays, r_hits = o.rays_view().copy(), o.hits_view().copy()
    finally:
        oracle.set_watertight(False, "32")
    # against the watertight restatement: only near ties may differ (walk order)
    mism = np.any(g_hits != r_hits, axis=-1)
    slots, ks = np.nonzero(mism)
    ties, _ = classify_mismatches(oracle, w.scene, w, slots, ks, g_hits[slots, ks], r_hits[slots, ks],
                                  g_rays[slots, ks, 3], edge=0.0)
    same = ~mism
    hit = same & (r_hits[..., 0] >= 0)
    et = rel_err(g_rays[..., 3][hit], r_rays[..., 3][hit], 0)
    e = rel_err(g_rays[..., :3][same], r_rays[..., :3][same])
    flipped = int((e > 1e-3).sum())
    print(f"{name}: vs watertight FP32 restatement: {mism.size} rays, id mismatches {int(mism.sum())} "
          f"(near ties {int(ties.sum())}), hit t {summary(et)}, radiance {summary(e)}, off by > 1e-3: {flipped}")
    assert int((~ties).sum()) == 0
    assert et.max() <= 1e-5
    assert np.percentile(e, 99.9) <= 1e-4 and flipped <= e.size // 10000
    # against Moller-Trumbore: differences only at edges and near ties
    d = np.any(g_hits != m_hits, axis=-1)
    slots, ks = np.nonzero(d)
    ties, grazes = classify_mismatches(oracle, w.scene, w, slots, ks, g_hits[slots, ks], m_hits[slots, ks],
                                       g_rays[slots, ks, 3], edge=1e-5)
    print(f"{name}: watertight vs Moller-Trumbore: {int(d.sum())} of {d.size} rays differ "
          f"(near ties {int(ties.sum())}, edge rays {int(grazes.sum())})")
    assert int((~(ties | grazes)).sum()) == 0


def test_watertight_closed_mesh_never_leaks_on_the_gpu():
    """Rays from inside a closed UV sphere aimed exactly at every vertex and every shared
    edge's f32 midpoint: the watertight field (watertight triangle test and conservative box
    tests) hits the sphere with every ray and agrees with its FP32 restatement; the
    Moller-Trumbore field's leaks (triangle edges and box faces) equal its restatement's."""
    v, t = sg.uv_sphere(24, 48, radius=1.3)
    v, t = np.asarray(v, np.float32), np.asarray(t, np.uint32)
    scene = _mesh_scene(v, t)
    centre = np.array([0.013, -0.021, 0.007])
    dirs = _edge_and_vertex_rays(v, t, centre)
    ro = np.tile(centre, (len(dirs), 1))
    w = sg.room_boxes(res=(2, 2, 2), rays=32)
    leaks = {}
    for trav, wt in ((N.TRAV_WATERTIGHT, True), (N.TRAV_MT_ORDERED, False)):
        f = ProbeField((2, 2, 2), (-1.0, -1.0, -1.0), (2.0, 2.0, 2.0), rays=32, traversal=trav)
        f.set_scene(scene)
        print(f"case: closed sphere, {len(dirs)} edge and vertex rays, traversal {trav}", flush=True)
        try:
            tt, tri, inst, _ = f.query(ro, dirs)
        finally:
            f.close()
        oracle.set_watertight(wt, "32")
        try:  # the restatement's own two-level walk: its box tests cull like the GPU's
            ref = [oracle.closest(scene, centre, d, "32") for d in dirs]
        finally:
            oracle.set_watertight(False, "32")
        ref_tri = np.array([r[1] for r in ref])
        leaks[trav] = (int((tri < 0).sum()), int((ref_tri < 0).sum()))
        assert np.array_equal(tri < 0, ref_tri < 0)
    del w
    print(f"leaks of {len(dirs)} rays: watertight {leaks[N.TRAV_WATERTIGHT]}, "
          f"Moller-Trumbore {leaks[N.TRAV_MT_ORDERED]} (GPU, restatement)")
    assert leaks[N.TRAV_WATERTIGHT] == (0, 0)
