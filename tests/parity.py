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
