"""Batched solves on the GPU (sbgi_batch_*): uneven fields — own scenes, cascades, ray
counts, blend precisions and traversal modes — updated by one batch frame must equal the
same fields updated one cascade at a time by sbgi_tick, bit for bit, on the eager first
frame, the captured frame and every hipGraph replay (replays interleaved with the
reference fields' own eager work and synchronisations).
"""
from __future__ import annotations

import numpy as np
import pytest

from speedball_amd import _native as N
from speedball_amd import scenegen as sg
from speedball_amd.field import Batch, ProbeField

pytestmark = pytest.mark.gpu


def workloads():
    """Three uneven fields: a 1-cascade room at 32 rays (BF16x3, ordered), a 2-cascade room
    with 100 boxes at 64 rays (FP32, compat), and a 3-cascade instanced scene at 128 rays
    (FP16, ordered, depth 16x16)."""
    a = sg.room_boxes(res=(6, 6, 6), rays=32, seed_id=11)
    b = sg.room_boxes(res=(8, 4, 6), rays=64, n_boxes=100, seed_id=12)
    c = sg.instanced(res=(8, 4, 6), rays=128, total_tris=24576, n_instances=32, n_lights=4)
    out = []
    for w, n_casc, prec, trav in ((a, 1, N.BLEND_BF16X3, N.TRAV_MT_ORDERED), (b, 2, N.BLEND_FP32, N.TRAV_MT_COMPAT),
                                  (c, 3, N.BLEND_FP16, N.TRAV_MT_ORDERED)):
        gm, gs = np.array(w.grid_min), np.array(w.grid_size)
        centre = gm + gs / 2
        casc = []
        for k in range(n_casc):
            size = gs / (2 ** (n_casc - 1 - k))
            casc.append((tuple((centre - size / 2).tolist()), tuple(size.tolist()), w.res))
        out.append((w, casc, prec, trav))
    return out


def make_fields(specs):
    fs = []
    for w, casc, prec, trav in specs:
        f = ProbeField(casc[0][2], casc[0][0], casc[0][1], rays=w.rays, depth_oct=w.depth_oct, precision=prec,
                       cascades=casc, traversal=trav)
        f.set_scene(w.scene)
        fs.append(f)
    return fs


def state(f):
    out = []
    for c in range(len(f.cascades)):
        out += [f.read_irr(c), f.read_depth(c), f.read_atlas(cascade=c), f.read_atlas(depth=True, cascade=c)]
    return out


def tick_individually(fields, params):
    for f, kw in zip(fields, params):
        for c in range(len(f.cascades)):
            f.tick(cascade=c, **kw)


@pytest.fixture(scope="module")
def specs():
    return workloads()


def test_batch_frames_equal_individual_ticks(specs):
    params = [dict(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis) for w, *_ in specs]
    ref, bat = make_fields(specs), make_fields(specs)
    batch = Batch(bat)
    try:
        flags = []
        for frame in range(5):
            if frame == 4:  # new parameters: eager frame, then a fresh capture on the next one
                params = [dict(kw, hysteresis=0.9) for kw in params]
            print(f"case: batch frame {frame}", flush=True)
            tick_individually(ref, params)
            batch.tick(params)
            ms, replayed = batch.timing()
            flags.append(replayed)
            print(f"frame {frame}: {ms:.3f} ms, replayed {replayed}")
            for i, (fr, fb) in enumerate(zip(ref, bat)):
                for a, b in zip(state(fr), state(fb)):
                    assert np.array_equal(a, b), (frame, i)
        assert flags == [False, False, True, True, False]
    finally:
        batch.close()
        for f in ref + bat:
            f.close()


@pytest.mark.parametrize("block", [512, 256])
def test_one_field_batch_equals_individual_ticks(specs, block):
    """A batch of one field (the 3-cascade instanced field): every segment walks the same
    scene, so the batched walks stage its TLAS and instance records in LDS; the frames must
    still equal the field's per-cascade ticks bit for bit, eager, captured and replayed."""
    one = [specs[2]]
    params = [dict(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis) for w, *_ in one]
    ref, bat = make_fields(one), make_fields(one)
    batch = Batch(bat)
    batch.set_walk_block(block)
    try:
        flags = []
        for frame in range(4):
            print(f"case: one-field batch, block {block}, frame {frame}", flush=True)
            tick_individually(ref, params)
            batch.tick(params)
            flags.append(batch.timing()[1])
            for a, b in zip(state(ref[0]), state(bat[0])):
                assert np.array_equal(a, b), frame
        assert flags == [False, False, True, True]
    finally:
        batch.close()
        for f in ref + bat:
            f.close()


def test_batch_autotune_keeps_results(specs):
    """The autotuner's schedule (walk tuning, block size, resident waves), applied to a
    fresh batch, must leave every frame equal to the individually ticked twin fields."""
    params = [dict(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis) for w, *_ in specs]
    tuned = make_fields(specs)
    batch = Batch(tuned)
    try:
        batch.tick(params)
        print("case: batch autotune", flush=True)
        tuning, block, ms = batch.autotune(params)
        print("batch autotune:", tuning, block, f"{ms:.3f} ms")
        schedule = {k: v for k, v in tuning.items() if k not in ("shadow_lists", "shadow_copies")}
        assert block in (256, 512) and ms > 0 and all(v >= 1 for v in schedule.values())
        assert tuning["shadow_lists"] == 0  # batch frames walk the TLAS
        assert all({k: f.walk_tuning()[k] for k in schedule} == schedule for f in tuned)
        waves = batch.walk_waves
        assert waves in (0, 8, 16)
        with pytest.raises(RuntimeError):
            batch.walk_waves = 2
    finally:
        batch.close()
        for f in tuned:
            f.close()
    # a schedule far from the defaults as well
    for tun, blk, wv in ((tuning, block, waves), (dict(tri_ratio=1, inst_ratio=16, refill_min=32, fetch_chunk=8), 256, 8)):
        ref, bat = make_fields(specs), make_fields(specs)
        for f in bat:
            f.set_walk_tuning(**tun)
        batch = Batch(bat)
        batch.set_walk_block(blk)
        batch.walk_waves = wv
        try:
            for frame in range(3):
                print(f"case: tuned batch {tun} block {blk} waves {wv} frame {frame}", flush=True)
                tick_individually(ref, params)
                batch.tick(params)
                for fr, fb in zip(ref, bat):
                    for a, b in zip(state(fr), state(fb)):
                        assert np.array_equal(a, b), frame
        finally:
            batch.close()
            for f in ref + bat:
                f.close()


def test_batch_rejects_bad_groups(specs):
    fs = make_fields(specs[:1])
    try:
        with pytest.raises(RuntimeError):
            Batch([fs[0], fs[0]])
        w, casc, prec, trav = specs[0]
        bare = ProbeField(casc[0][2], casc[0][0], casc[0][1], rays=w.rays, depth_oct=w.depth_oct)
        try:
            with pytest.raises(RuntimeError):
                Batch([fs[0], bare])  # no scene
        finally:
            bare.close()
    finally:
        for f in fs:
            f.close()


def test_batch_refuses_a_field_whose_scene_outgrew_it(specs):
    """The batch sizes each segment's shadow queue by the field's light count at creation:
    after set_scene with more lights, a frame is refused instead of dropping records."""
    fs = make_fields(specs[:1])
    w = specs[0][0]
    params = [dict(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)]
    batch = Batch(fs)
    try:
        batch.tick(params)
        bigger = sg.instanced(res=(4, 4, 4), rays=32, total_tris=6000, n_instances=8, n_lights=4).scene
        fs[0].set_scene(bigger)
        with pytest.raises(RuntimeError):
            batch.tick(params)
    finally:
        batch.close()
        for f in fs:
            f.close()


def test_batch_frames_match_the_oracle(specs):
    """Two batch frames (from cleared state, then bouncing off the first) against the CPU
    oracle, per field and cascade (each cascade solve reads only its own atlas): the FP32
    field against the FP32 restatement at the FP32 bars (p99 1e-4, max 1e-3), the MFMA
    fields against the FP64 oracle fed the reference's f32 ray directions at the blend bars
    (p99.9 1e-3; BF16x3 and FP16), over both chained updates."""
    import oracle
    from tests.parity import interior_irr, rel_err, summary
    params = [dict(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis) for w, *_ in specs]
    fields = make_fields(specs)
    batch = Batch(fields)
    try:
        got = []
        for frame in range(2):
            print(f"case: batch frame {frame} for the oracle comparison", flush=True)
            batch.tick(params)
            batch.wait()
            got.append([[f.read_irr(c) for c in range(len(f.cascades))] for f in fields])
    finally:
        batch.close()
        for f in fields:
            f.close()
    for i, ((w, casc, prec, trav), kw) in enumerate(zip(specs, params)):
        fp32 = prec == N.BLEND_FP32
        for c, (gm, gs, res) in enumerate(casc):
            oracle.set_f32_dirs(not fp32)
            try:
                o = oracle.OracleField(w.scene, gm, gs, res, w.rays, w.depth_oct, prec="32" if fp32 else "64")
                p = o.params(**kw)
                for frame in range(2):
                    o.update(p)
                    e = rel_err(interior_irr(got[frame][i][c])[..., :3], interior_irr(o.irr_view())[..., :3])
                    print(f"field {i} (precision {prec}) cascade {c} frame {frame}: irradiance vs "
                          f"{'FP32' if fp32 else 'FP64'} {summary(e)}")
                    if fp32:
                        assert np.percentile(e, 99) <= 1e-4 and e.max() <= 1e-3
                    else:
                        assert np.percentile(e, 99.9) <= 1e-3
            finally:
                oracle.set_f32_dirs(False)
