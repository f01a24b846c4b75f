"""Shared parity metrics for the GPU tests (imported by tests only)."""
from __future__ import annotations

import numpy as np


def rel_err(a, b, floor_frac=1e-2):
    """|a-b| / max(|b|, floor) with floor = floor_frac * mean |b|: texels that are black or
    near-black in the reference (below 1% of the field mean) are measured against that floor,
    where a relative error is otherwise undefined."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    floor = max(floor_frac * float(np.mean(np.abs(b))), 1e-30)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def summary(e):
    e = np.asarray(e).reshape(-1)
    return {"max": float(e.max()) if e.size else 0.0, "p99": float(np.percentile(e, 99)) if e.size else 0.0,
            "mean": float(e.mean()) if e.size else 0.0}


def interior_irr(irr):
    """(P, 64, 4) irradiance state -> (P, 36, 4) interior texels."""
    idx = [(ly + 1) * 8 + (lx + 1) for ly in range(6) for lx in range(6)]
    return irr[:, idx, :]


def near_tie_mask(oracle_mod, scene, field_w, slots_probes, ks, t_gpu=None, tol=1e-6):
    """For each mismatched (probe, k) ray, True when its two nearest hits lie within tol
    relative: either the FP64 brute force sees a second accepted hit within tol of the
    first, or the GPU's (different) hit is at the same distance as the oracle's within tol
    (an adjacent triangle across a shared edge, accepted one ulp inside its barycentric
    range)."""
    out = []
    for n, (probe, k) in enumerate(zip(slots_probes, ks)):
        ro = probe_pos(field_w, probe)
        rd = oracle_mod.ray_dir(int(k), field_w.rays, field_w.frame_jitter)
        t, tri, inst, t2 = oracle_mod.brute_closest(scene, ro, rd)
        tie = tri >= 0 and t2 < 1e29 and abs(t2 - t) <= tol * abs(t)
        if not tie and t_gpu is not None and tri >= 0 and t_gpu[n] >= 0:
            tie = abs(t_gpu[n] - t) <= tol * max(abs(t), 1.0)
        out.append(tie)
    return np.array(out, bool)


def probe_pos(w, p):
    rx, ry, rz = w.res
    ix, iy, iz = p % rx, (p // rx) % ry, p // (rx * ry)
    f = np.array([ix / max(1, rx - 1), iy / max(1, ry - 1), iz / max(1, rz - 1)], np.float64)
    return np.asarray(w.grid_min, np.float64) + f * np.asarray(w.grid_size, np.float64)


def edge_distance(scene, ro, rd, tri, inst):
    """FP64 barycentric distance of the ray to the nearest edge of triangle `tri` of
    instance slot `inst`: min(u, v, 1 - u - v) of the ray/plane intersection in the
    instance's local frame (negative when the ray passes outside)."""
    if tri < 0:
        return np.inf
    rows = scene.materials[scene.inst_base + inst * 28: scene.inst_base + inst * 28 + 12].astype(np.float64)
    A = rows.reshape(3, 4)
    lro = A[:, :3] @ np.asarray(ro, np.float64) + A[:, 3]
    lrd = A[:, :3] @ np.asarray(rd, np.float64)
    v = scene.vertex_data.reshape(-1, 8)[:, :3].astype(np.float64)
    i0, i1, i2 = scene.tri_index.reshape(-1, 3)[tri]
    p0, e1, e2 = v[i0], v[i1] - v[i0], v[i2] - v[i0]
    pv = np.cross(lrd, e2)
    det = e1 @ pv
    if abs(det) < 1e-300:
        return 0.0
    tv = lro - p0
    u = (tv @ pv) / det
    qv = np.cross(tv, e1)
    w = (lrd @ qv) / det
    return min(u, w, 1.0 - u - w)


def classify_mismatches(oracle_mod, scene, field_w, slots, ks, gpu_hits, ref_hits, t_gpu, tol=1e-6, edge=1e-5):
    """Split hit-id mismatches into near ties, edge grazes (the ray passes within `edge`
    barycentric units of an edge of either triangle: Moller-Trumbore's edge decisions there
    depend on the last bit of the arithmetic) and real disagreements."""
    ties = near_tie_mask(oracle_mod, scene, field_w, slots, ks, t_gpu, tol)
    grazes = np.zeros(len(slots), bool)
    for n, (probe, k) in enumerate(zip(slots, ks)):
        if ties[n]:
            continue
        ro = probe_pos(field_w, probe)
        rd = oracle_mod.ray_dir(int(k), field_w.rays, field_w.frame_jitter)
        dg = edge_distance(scene, ro, rd, gpu_hits[n][0], gpu_hits[n][1])
        dr = edge_distance(scene, ro, rd, ref_hits[n][0], ref_hits[n][1])
        grazes[n] = abs(dg) < edge or abs(dr) < edge
    return ties, grazes
