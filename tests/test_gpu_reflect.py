"""Reflection lobes on the GPU (rough power-8 lobe, glossy power-64 cache; -m gpu).

Kernel parity is checked in isolation: the oracle's lobes are evaluated on the GPU's own
ray records, depth moments and history of the same update, so the comparison covers the
contraction, the emitter injection and the temporal policy alone. Bars: FP32 lobes within
1e-5 relative of the FP32 restatement (max; floor 1e-3 of the mean); MFMA lobes within
1e-3 (BF16x3, FP16) and 1e-2 (BF16) of FP64 at p99.9; atlases equal to the states they
store, rounded to FP16; batch frames equal per-field ticks bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from speedball_amd import _native as N
from speedball_amd import scenegen as sg
from speedball_amd.field import Batch, ProbeField
from tests.parity import rel_err, summary

pytestmark = pytest.mark.gpu


def metal_room():
    """The room of room_boxes with a third of the boxes metal (unshaded hits)."""
    w = sg.room_boxes(res=(6, 6, 6), rays=64)
    s = w.scene
    mats = s.materials.copy()
    n_mat = s.inst_base // 28
    for m in range(1, n_mat, 3):
        mats[m * 28 + 4] = 1.0  # metalness
    s2 = sg.Scene(**{k: getattr(s, k) for k in ("bvh_nodes", "tri_index", "vertex_data", "tri_material", "lights",
                                                "inst_base", "inst_count", "tlas_base", "tlas_node_count",
                                                "tlas_order")}, materials=mats)
    w.scene = s2
    return w


def make(w, q, precision=N.BLEND_FP32, depth_oct=6):
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=depth_oct, precision=precision,
                   traversal=N.TRAV_MT_ORDERED, reflections=q)
    f.set_scene(w.scene)
    return f


def params(w, phase=0, **kw):
    return dict(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis, sky_intensity=1.0, sky_configured=1.0,
                reflection_sky_fallback=1.0, glossy_phase=phase, sky_sh=[0.5] * 3 + [0.0] * 24, **kw)


def lobes(f, G):
    P = f.probes()
    out = {"rough": f.read(N.BUF_ROUGH_STATE, np.float32).reshape(P, 64, 4)}
    if G:
        out["gnum"] = f.read(N.BUF_GLOSSY_STATE, np.float32).reshape(P, G * G, 4)
        out["gden"] = f.read(N.BUF_GLOSSY_WEIGHT, np.float32).reshape(P, G * G)
    return out


def isolated_step(f, w, q, prec_ref, kw):
    """One GPU update, and the oracle's lobes of that update from the GPU's own inputs."""
    G = {2: 10, 3: 18}.get(q, 0)
    before = lobes(f, G)
    depth_before = f.read(N.BUF_DEPTH_STATE, np.float32)
    f.tick(**kw)
    after = lobes(f, G)
    o = oracle.OracleField(w.scene, w.grid_min, w.grid_size, w.res, w.rays, f.depth_oct, prec=prec_ref,
                           reflections=q)
    dt = o.dt
    o.ray_data = f.read_rays().astype(dt).reshape(-1)
    o.depth[:] = depth_before.astype(dt)
    o.rough[:] = before["rough"].astype(dt).reshape(-1)
    o.evis[:] = f.read_emitter_vis().astype(dt).reshape(-1)
    o.meta[:] = f.read(N.BUF_META, np.float32)
    p = o.params(**kw)
    o.reflect_rough(p)
    ref = {"rough": o.rough.reshape(-1, 64, 4)}
    if G:
        o.gnum[:] = before["gnum"].astype(dt).reshape(-1)
        o.gden[:] = before["gden"].astype(dt).reshape(-1)
        o.reflect_glossy(p)
        ref["gnum"] = o.gnum.reshape(-1, G * G, 4)
        ref["gden"] = o.gden.reshape(-1, G * G)
    return after, ref


INTERIOR = [(t // 6 + 1) * 8 + t % 6 + 1 for t in range(36)]


def compare(after, ref, bar, pct=None, label=""):
    for k in after:
        g = after[k][:, INTERIOR] if k == "rough" else after[k]
        r = ref[k][:, INTERIOR] if k == "rough" else ref[k]
        e = rel_err(g, r, floor_frac=1e-3)
        print(f"{label} {k}: {summary(e)}")
        if pct is None:
            assert e.max() <= bar, k
        else:
            assert np.percentile(e, pct) <= bar, k


@pytest.mark.parametrize("q", [N.REFLECT_ROUGH, N.REFLECT_HIGH, N.REFLECT_ULTRA], ids=["rough", "high", "ultra"])
def test_fp32_lobes_equal_the_restatement(q):
    w = metal_room()
    f = make(w, q)
    try:
        for step in range(3):
            print(f"case: fp32 lobes q {q} step {step}", flush=True)
            after, ref = isolated_step(f, w, q, "32", params(w, phase=step % 2))
            compare(after, ref, 1e-5, label=f"step {step}")
        rays = f.read_rays()
        assert (rays[..., 3] < -1.5).sum() > 100  # metal hits carry the unshaded code
        # the rough atlas stores the rough state, FP16, with mirrored gutters
        ra = f.read(N.BUF_ROUGH_ATLAS, np.uint16).view(np.float16).astype(np.float32)
        ra = ra.reshape(w.res[1] * w.res[2] * 8, w.res[0] * 8, 4)
        st = after["rough"]
        probe = 5
        col, row = probe % w.res[0], (probe // (w.res[0] * w.res[1])) * w.res[1] + (probe // w.res[0]) % w.res[1]
        tile = ra[row * 8:(row + 1) * 8, col * 8:(col + 1) * 8]
        np.testing.assert_array_equal(tile[1:7, 1:7].reshape(36, 4), st[probe, INTERIOR].astype(np.float16))
        np.testing.assert_array_equal(tile[0, 3], tile[1, 4])  # top gutter mirrors (edge - lx, 1)
        if q != N.REFLECT_ROUGH:
            G = 18 if q == N.REFLECT_ULTRA else 10
            ga = f.read(N.BUF_GLOSSY_ATLAS, np.uint16).view(np.float16).astype(np.float32)
            tx = int(np.ceil(np.sqrt(f.probes())))
            ga = ga.reshape(-1, tx * G, 4)
            num, den = after["gnum"][probe], after["gden"][probe]
            gt = ga[(probe // tx) * G:(probe // tx + 1) * G, (probe % tx) * G:(probe % tx + 1) * G].reshape(G * G, 4)
            want = np.concatenate([num[:, :3] / den[:, None], np.clip(num[:, 3:] / den[:, None], 0, 1)], 1)
            solved = den > 0
            np.testing.assert_array_equal(gt[solved], want[solved].astype(np.float16))
    finally:
        f.close()


@pytest.mark.parametrize("precision,bar", [(N.BLEND_BF16X3, 1e-3), (N.BLEND_FP16, 1e-3), (N.BLEND_BF16, 1e-2)],
                         ids=["bf16x3", "fp16", "bf16"])
def test_mfma_lobes_within_their_bars(precision, bar):
    w = metal_room()
    f = make(w, N.REFLECT_ULTRA, precision=precision, depth_oct=14)
    try:
        for step in range(2):
            print(f"case: mfma lobes precision {precision} step {step}", flush=True)
            after, ref = isolated_step(f, w, N.REFLECT_ULTRA, "64", params(w))
            compare(after, ref, bar, pct=99.9, label=f"precision {precision} step {step}")
    finally:
        f.close()


def test_lobes_with_emitters():
    w = sg.room_lights(res=(6, 6, 6), rays=64, n_point=8, n_emitters=4)
    f = make(w, N.REFLECT_ULTRA)
    try:
        for step in range(2):
            print(f"case: lobes with emitters step {step}", flush=True)
            after, ref = isolated_step(f, w, N.REFLECT_ULTRA, "32", params(w, emitter_vis_seed=0.3))
            compare(after, ref, 1e-5, label=f"emitters step {step}")
    finally:
        f.close()


def test_lobes_end_to_end_against_the_oracle_update():
    """The whole update with lobes against the FP32 oracle's own chained updates
    (primary-ray near ties aside, the lobes agree at the parity bar)."""
    w = metal_room()
    f = make(w, N.REFLECT_ULTRA)
    o = oracle.OracleField(w.scene, w.grid_min, w.grid_size, w.res, w.rays, 6, prec="32", reflections=3)
    try:
        kw = params(w)
        for step in range(2):
            print(f"case: lobes end to end step {step}", flush=True)
            f.tick(**kw)
            o.update(o.params(**kw))
        g = lobes(f, 18)
        e = rel_err(g["rough"][:, INTERIOR], o.rough.reshape(-1, 64, 4)[:, INTERIOR], floor_frac=1e-3)
        eg = rel_err(g["gnum"], o.gnum.reshape(-1, 324, 4), floor_frac=1e-3)
        print("end to end rough", summary(e), "glossy numerator", summary(eg))
        assert np.percentile(e, 99) <= 1e-4 and np.percentile(eg, 99) <= 1e-4
    finally:
        f.close()


def test_batch_with_lobes():
    w = metal_room()
    ref, bat = [make(w, N.REFLECT_HIGH, precision=N.BLEND_BF16X3)], [make(w, N.REFLECT_HIGH, precision=N.BLEND_BF16X3)]
    batch = Batch(bat)
    try:
        for frame in range(4):
            print(f"case: lobes batch frame {frame}", flush=True)
            kw = params(w, phase=frame % 2)
            ref[0].tick(**kw)
            batch.tick([kw])
            for k in (N.BUF_ROUGH_STATE, N.BUF_GLOSSY_STATE, N.BUF_GLOSSY_WEIGHT, N.BUF_GLOSSY_ATLAS,
                      N.BUF_ROUGH_ATLAS, N.BUF_IRR_STATE):
                assert np.array_equal(ref[0].read(k, np.uint32), bat[0].read(k, np.uint32)), (frame, k)
    finally:
        batch.close()
        ref[0].close()
        bat[0].close()
