"""Clustered light grid and emitter lights on the GPU (run on an MI355X with -m gpu).

Bars: the cell lists equal the oracle's word for word; a clustered field equals the same
field without clustering bit for bit (the lists only skip lights that cannot reach the
hit); retained emitter visibility equals the FP32 restatement except rays whose any-hit
answer flips on the hardware reciprocal (at most 1 in 1,000 probe-emitter pairs, counted);
irradiance with emitters within 1e-4 relative of the FP32 restatement (p99) and 1e-3 of the
FP64 oracle (p99.9); a batch frame with emitters and clustering equals individual ticks
bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from speedball_amd import _native as N
from speedball_amd import scenegen as sg
from speedball_amd.field import Batch, ProbeField
from tests.parity import interior_irr, rel_err, summary

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lights_room():
    return sg.room_lights(res=(8, 8, 8), rays=64, n_point=24, n_emitters=4)


def make(w, clustered=False, precision=N.BLEND_FP32, traversal=N.TRAV_MT_ORDERED):
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=w.depth_oct, precision=precision,
                   keep_hit_ids=True, traversal=traversal, clustered=clustered)
    f.set_scene(w.scene)
    return f


def params(w, **kw):
    return dict(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis, emitter_vis_seed=0.41, **kw)


def test_light_grid_equals_oracle(lights_room):
    w = lights_room
    print("case: light grid readback", flush=True)
    f = make(w, clustered=True)
    try:
        got = f.read_light_grid()
        want = oracle.light_grid(w.scene, oracle.cascade_desc(w.grid_min, w.grid_size, w.res)).reshape(-1, 17)
        assert np.array_equal(got, want)
        # new light records rebuild the lists: move every ranged point light by +1 m in x
        lights = w.scene.lights.copy().reshape(-1, 17)
        lights[lights[:, 0] == 1, 1] += 1.0
        f.update_ranges(N.BUF_LIGHTS, [(0, lights.size)], lights.reshape(-1))
        s2 = sg.Scene(**{k: getattr(w.scene, k) for k in ("bvh_nodes", "tri_index", "vertex_data", "tri_material",
                                                           "materials", "inst_base", "inst_count", "tlas_base",
                                                           "tlas_node_count", "tlas_order", "emitter_base",
                                                           "emitter_count")},
                      lights=lights.reshape(-1))
        want2 = oracle.light_grid(s2, oracle.cascade_desc(w.grid_min, w.grid_size, w.res)).reshape(-1, 17)
        assert not np.array_equal(want, want2)
        assert np.array_equal(f.read_light_grid(), want2)
    finally:
        f.close()


@pytest.mark.parametrize("precision", [N.BLEND_FP32, N.BLEND_BF16X3], ids=["fp32", "bf16x3"])
def test_clustered_equals_all_lights(lights_room, precision):
    w = lights_room
    a, b = make(w, precision=precision), make(w, clustered=True, precision=precision)
    try:
        for step in range(3):
            print(f"case: clustered vs all lights, step {step}", flush=True)
            a.tick(**params(w))
            b.tick(**params(w))
            assert np.array_equal(a.read_rays(), b.read_rays()), step
            assert np.array_equal(a.read_irr(), b.read_irr()), step
            assert np.array_equal(a.read_atlas(), b.read_atlas()), step
            assert np.array_equal(a.read_emitter_vis(), b.read_emitter_vis()), step
    finally:
        a.close()
        b.close()


@pytest.fixture(scope="module")
def emitter_run(lights_room):
    """Three chained updates with emitters (FP32 blend, clustered) on the GPU and in the FP32
    restatement and FP64 oracle."""
    w = lights_room
    f = make(w, clustered=True)
    gpu = []
    for step in range(3):
        print(f"case: emitter field, step {step}", flush=True)
        f.tick(**params(w))
        gpu.append(dict(evis=f.read_emitter_vis(), irr=f.read_irr(), hits=f.read_hit_ids()))
    f.close()
    refs = {}
    for prec in ("32", "64"):
        o = oracle.OracleField(w.scene, w.grid_min, w.grid_size, w.res, w.rays, w.depth_oct, prec=prec)
        o.set_clustered()
        p = o.params(**params(w))
        steps = []
        for _ in range(3):
            o.update(p)
            steps.append(dict(evis=o.evis.reshape(-1, 16).copy(), irr=o.irr_view().copy(),
                              hits=o.hits_view().copy()))
        refs[prec] = steps
    return w, gpu, refs


def test_emitter_visibility(emitter_run):
    w, gpu, refs = emitter_run
    E = w.scene.emitter_count
    for step in range(3):
        g, r = gpu[step]["evis"], refs["32"][step]["evis"]
        assert np.all(g[:, E:] == 0)
        off = np.abs(g[:, :E] - r[:, :E]) > 1e-6
        print(f"\nstep {step}: emitter visibility values {np.unique(np.round(g[:, :E], 4))}, "
              f"pairs differing from FP32 {int(off.sum())} of {off.size}")
        assert off.sum() <= max(1, off.size // 1000)
        assert np.array_equal(g[:, :E][~off], r[:, :E][~off])
    assert (gpu[2]["evis"][:, :E] > 0.4).mean() > 0.2  # most probes see some emitter


def test_emitter_irradiance(emitter_run):
    w, gpu, refs = emitter_run
    for step in range(3):
        g = interior_irr(gpu[step]["irr"])
        e32 = rel_err(g[..., :3], interior_irr(refs["32"][step]["irr"])[..., :3])
        e64 = rel_err(g[..., :3], interior_irr(refs["64"][step]["irr"])[..., :3])
        print(f"\nstep {step} irradiance with emitters vs FP32 {summary(e32)} vs FP64 {summary(e64)}")
        assert np.percentile(e32, 99) <= 1e-4
        assert np.percentile(e64, 99.9) <= 1e-3


def test_emitter_blend_mfma_precisions(lights_room):
    """The MFMA blends inject the same virtual rays: BF16x3 / FP16 / BF16 within their
    stated bars of the FP64 oracle after one update."""
    w = lights_room
    o = oracle.OracleField(w.scene, w.grid_min, w.grid_size, w.res, w.rays, w.depth_oct, prec="64")
    o.update(o.params(**params(w)))
    r64 = interior_irr(o.irr_view())[..., :3]
    for prec, bar in ((N.BLEND_BF16X3, 1e-3), (N.BLEND_FP16, 1e-3), (N.BLEND_BF16, 1e-2)):
        print(f"case: emitter blend precision {prec}", flush=True)
        f = make(w, precision=prec)
        try:
            f.tick(**params(w))
            e = rel_err(interior_irr(f.read_irr())[..., :3], r64)
        finally:
            f.close()
        print(f"precision {prec}: irradiance vs FP64 {summary(e)}")
        assert np.percentile(e, 99) <= bar


def test_emitter_visibility_clears_with_the_atlas(lights_room):
    w = lights_room
    f = make(w)
    try:
        print("case: emitter visibility clear", flush=True)
        f.tick(**params(w))
        f.tick(**params(w))
        before = f.read_emitter_vis()
        assert before.max() > 0.3
        f.needs_clear[0] = True
        f.tick(**params(w))
        after = f.read_emitter_vis()
        assert after.max() <= 0.2 + 1e-6  # one sample after the clear: 0 or 0.2
        f.write(N.BUF_EMITTER_VIS, before)
        assert np.array_equal(f.read_emitter_vis(), before)
    finally:
        f.close()


def test_batch_with_emitters_and_clustering(lights_room):
    """A batch frame over a clustered emitter field and a plain field equals individual
    ticks bit for bit, on the eager, captured and replayed frames."""
    w = lights_room
    w2 = sg.room_boxes(res=(6, 6, 6), rays=32, seed_id=11)

    def fields():
        a = make(w, clustered=True, precision=N.BLEND_BF16X3)
        b = ProbeField(w2.res, w2.grid_min, w2.grid_size, rays=w2.rays, depth_oct=w2.depth_oct,
                       traversal=N.TRAV_MT_ORDERED)
        b.set_scene(w2.scene)
        return [a, b]

    ps = [params(w), dict(frame_jitter=w2.frame_jitter, hysteresis=w2.hysteresis)]
    ref, bat = fields(), fields()
    batch = Batch(bat)
    try:
        for frame in range(4):
            print(f"case: emitter batch frame {frame}", flush=True)
            for f, kw in zip(ref, ps):
                f.tick(**kw)
            batch.tick(ps)
            for fr, fb in zip(ref, bat):
                assert np.array_equal(fr.read_irr(), fb.read_irr()), frame
                assert np.array_equal(fr.read_emitter_vis(), fb.read_emitter_vis()), frame
                assert np.array_equal(fr.read_atlas(), fb.read_atlas()), frame
    finally:
        batch.close()
        for f in ref + bat:
            f.close()


def test_bad_emitter_and_cluster_descriptions(lights_room):
    w = lights_room
    with pytest.raises(RuntimeError):
        ProbeField((1, 4, 4), w.grid_min, w.grid_size, rays=32, clustered=True)  # needs res >= 2
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=32)
    try:
        v = w.scene.view()
        old = (v.emitter_base, v.emitter_count)
        try:
            v.emitter_count = 17
            with pytest.raises(RuntimeError):
                f.set_scene(w.scene)
            v.emitter_count = old[1]
            v.emitter_base = v.light_count
            with pytest.raises(RuntimeError):
                f.set_scene(w.scene)
        finally:
            v.emitter_base, v.emitter_count = old
        f.set_scene(w.scene)
    finally:
        f.close()
