"""Differentiable mode on the GPU: forward, hand-written adjoint, finite differences.

The differentiable solve (include/sbgi.h, sbgi_diff_*) is two first-updates from
cleared state, the second bouncing off the first's unrounded FP32 atlas. Bars:
- forward: interior irradiance within 1e-4 relative (p99) of the FP64 oracle's;
- adjoint: gradients of sum (E2 - T)^2 with respect to per-instance-triangle albedo and
  light intensities within 1e-3 of the FP64 oracle adjoint (relative to the largest
  gradient entry) and within 1e-3 relative of FP64 central differences (eps 1e-3) on 64
  seeded parameters plus every light;
- forward + adjoint <= 3 x forward at the full differentiable workload (4 tori x 1024
  triangles x 64 instances, 16 lights, 16^3 probes x 128 rays).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from speedball_amd import scenegen as sg
from speedball_amd.field import ProbeField

pytestmark = pytest.mark.gpu

SMALL = dict(res=(4, 4, 4), rays=64, n_mesh=2, n_instances=6, torus_n=4, n_lights=3, box=8.0)


def _field(w):
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=6, precision=0, traversal=1)
    f.set_scene(w.scene)
    return f


@pytest.fixture(scope="module")
def small():
    w, albedo, target_albedo = sg.differentiable(**SMALL)
    albedo = albedo * 0.9
    f = _field(w)
    params = dict(frame_jitter=w.frame_jitter, hysteresis=0.0)
    p64 = oracle.OracleField(w.scene, w.grid_min, w.grid_size, w.res, w.rays, 6, prec="64").params(**params)
    yield w, f, albedo, params, p64
    f.close()


def test_albedo_map_matches_oracle(small):
    w, f, *_ = small
    m64, n = oracle.albedo_map(w.scene)
    assert np.array_equal(f.diff_albedo_map(), m64)
    assert f.diff_param_count() == (n, SMALL["n_lights"])


def test_forward_matches_fp64(small):
    w, f, albedo, params, p64 = small
    ls = np.array([1.0, 0.8, 1.2], np.float32)
    print(f"case: diff forward {w.res} x {w.rays}", flush=True)
    E2 = f.diff_forward(albedo, ls, **params)
    amap, _ = oracle.albedo_map(w.scene)
    E64, _, _ = oracle.diff_forward(w.scene, w.grid_min, w.grid_size, w.res, w.rays, p64, albedo.astype(np.float64),
                                    amap, ls.astype(np.float64))
    floor = 1e-2 * np.abs(E64).mean()
    e = np.abs(E2 - E64) / np.maximum(np.abs(E64), floor)
    print("forward rel err p99", np.percentile(e, 99), "max", e.max())
    assert np.percentile(e, 99) <= 1e-4


def test_adjoint_matches_fp64_adjoint_and_central_differences(small):
    import torch
    from speedball_amd.autograd import diff_probe_update
    w, f, albedo, params, p64 = small
    amap, n = oracle.albedo_map(w.scene)
    i0 = w.scene.lights.reshape(-1, 17)[:, 7].astype(np.float64)  # colour x intensity of white lights
    intensity = i0 * np.array([1.0, 0.8, 1.2])
    a_t = torch.tensor(albedo, dtype=torch.float32, requires_grad=True)
    i_t = torch.tensor(intensity, dtype=torch.float32, requires_grad=True)
    print("case: adjoint vs fp64", flush=True)
    E2 = diff_probe_update(a_t, i_t, f, i0, params)
    rng = np.random.default_rng(3)
    T = torch.tensor(E2.detach().numpy() * (1 + rng.uniform(-0.3, 0.3, E2.shape)), dtype=torch.float32)
    loss = ((E2 - T) ** 2).sum()
    loss.backward()
    ga = a_t.grad.numpy().astype(np.float64)
    gi = i_t.grad.numpy().astype(np.float64)
    # FP64 oracle adjoint at the same point (light scale = intensity / i0)
    ls = intensity / i0
    args = (w.scene, w.grid_min, w.grid_size, w.res, w.rays, p64)
    E64, _, A1 = oracle.diff_forward(*args, albedo.astype(np.float64), amap, ls)
    g64a, g64l = oracle.diff_adjoint(*args, albedo.astype(np.float64), amap, ls, A1, 2 * (E64 - T.numpy()))
    g64i = g64l / i0
    scale = np.abs(g64a).max()
    da = np.abs(ga - g64a)
    print("albedo grad: max |gpu - fp64| / max|g| =", da.max() / scale, " entries", n)
    assert da.max() <= 1e-3 * scale
    np.testing.assert_allclose(gi, g64i, rtol=1e-3, atol=1e-3 * np.abs(g64i).max())

    # central differences of the FP64 forward on 64 seeded albedo parameters (among the
    # entries a ray reaches) and on every light intensity
    def loss64(a, l):
        return float(((oracle.diff_forward(*args, a, amap, l)[0] - T.numpy().astype(np.float64)) ** 2).sum())

    eps = 1e-3
    live = np.flatnonzero(np.abs(g64a).reshape(-1) > 1e-6 * scale)
    pick = np.random.default_rng(0x5BA11004).choice(live, size=min(64, len(live)), replace=False)
    a64 = albedo.astype(np.float64)
    bad = 0
    for fidx in pick:
        i, c = divmod(int(fidx), 3)
        ap, am = a64.copy(), a64.copy()
        ap[i, c] += eps
        am[i, c] -= eps
        fd = (loss64(ap, ls) - loss64(am, ls)) / (2 * eps)
        if abs(fd - ga[i, c]) > 1e-3 * max(abs(fd), 1e-2 * scale):
            bad += 1
            print("  mismatch", i, c, fd, ga[i, c])
    assert bad == 0
    for j in range(len(ls)):
        lp, lm = ls.copy(), ls.copy()
        lp[j] += eps
        lm[j] -= eps
        fd = (loss64(a64, lp) - loss64(a64, lm)) / (2 * eps) / i0[j]
        assert abs(fd - gi[j]) <= 1e-3 * max(abs(fd), 1e-6), (j, fd, gi[j])


def test_forward_plus_adjoint_within_3x_forward():
    w, albedo, target = sg.differentiable()
    f = _field(w)
    params = dict(frame_jitter=w.frame_jitter, hysteresis=0.0)
    print(f"case: diff timing {w.res} x {w.rays}, {w.scene.tri_count} unique triangles", flush=True)
    E2 = f.diff_forward(albedo, **params)
    g = 2 * (E2 - f.diff_forward(target, **params))
    f.diff_forward(albedo, **params)  # the point the adjoint is taken at
    fwd, adj = [], []
    for _ in range(3):
        f.diff_forward(albedo, **params)
        f.diff_adjoint(g)
        a, b = f.diff_timing()
        fwd.append(a)
        adj.append(b)
    f.close()
    fwd, adj = float(np.median(fwd)), float(np.median(adj))
    print(f"forward {fwd:.3f} ms, adjoint {adj:.3f} ms, ratio (fwd+adj)/fwd = {(fwd + adj) / fwd:.2f}")
    assert fwd + adj <= 3 * fwd
