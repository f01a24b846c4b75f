"""Shared parity metrics for the GPU tests (imported by tests only)."""
from __future__ import annotations

import numpy as np


def rel_err(a, b, floor_frac=1e-3):
    """|a-b| / max(|b|, floor) with floor = floor_frac * mean |b| (texels that are exactly
    black in the reference would otherwise make a relative error undefined)."""
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


def near_tie_mask(oracle_mod, scene, field_w, slots_probes, ks, tol=1e-6):
    """For each (probe, k) ray, True when the two nearest accepted hits are within tol relative."""
    out = []
    for probe, k in zip(slots_probes, ks):
        ro = probe_pos(field_w, probe)
        rd = oracle_mod.ray_dir(int(k), field_w.rays, field_w.frame_jitter)
        t, tri, inst, t2 = oracle_mod.brute_closest(scene, ro, rd)
        out.append(tri >= 0 and t2 < 1e29 and abs(t2 - t) <= tol * abs(t))
    return np.array(out, bool)


def probe_pos(w, p):
    rx, ry, rz = w.res
    ix, iy, iz = p % rx, (p // rx) % ry, p // (rx * ry)
    f = np.array([ix / max(1, rx - 1), iy / max(1, ry - 1), iz / max(1, rz - 1)], np.float64)
    return np.asarray(w.grid_min, np.float64) + f * np.asarray(w.grid_size, np.float64)
