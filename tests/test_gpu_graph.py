"""The tick's hipGraph (run on an MI355X with -m gpu).

A field that replays its captured tick graph and a twin created with SBGI_FIELD_NO_GRAPH
(every tick eager) must hold the same state bit for bit after every update, through
several replays with the program's own eager work and synchronisations in between (state
reads, ray queries that launch the walk eagerly on the same stream), a parameter change
that forces a recapture, and a scene edit between replays (device data changes, the
captured arguments do not). With stage timing on, the stage events of replayed ticks keep
timing their stages; off (the default), the graph has no event nodes and the stats say
that their stage times are not current.
"""
from __future__ import annotations

import numpy as np
import pytest

from speedball_amd import _native as N
from speedball_amd import scenegen as sg
from speedball_amd.field import ProbeField

pytestmark = pytest.mark.gpu


def _state(f):
    return [f.read_irr(), f.read_depth(), f.read_atlas(), f.read_atlas(depth=True), f.read_rays()]


@pytest.mark.parametrize("timing", [False, True], ids=["no-events", "stage-events"])
@pytest.mark.parametrize("precision,depth_oct", [(N.BLEND_BF16X3, 14), (N.BLEND_FP32, 6)])
def test_tick_graph_replays_equal_eager_ticks(precision, depth_oct, timing):
    w = sg.instanced(res=(8, 4, 6), rays=64, total_tris=24576, n_instances=32, n_lights=4)
    fields = [ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=depth_oct, precision=precision,
                         traversal=N.TRAV_MT_ORDERED, graph=g) for g in (True, False)]
    rng = np.random.default_rng(3)
    try:
        for f in fields:
            f.set_scene(w.scene)
        fields[0].set_stage_timing(timing)
        kw = dict(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
        replays = 0
        for step in range(9):
            if step == 5:
                kw = dict(frame_jitter=0.37, hysteresis=0.8)  # new arguments: recapture
            if step == 7:
                # move two instances: device data changes, the captured arguments do not
                slots = np.array([0, 5], np.uint32)
                worlds = [np.eye(4), np.eye(4)]
                for m in worlds:
                    m[:3, 3] = rng.uniform(2, 20, 3)
                for f in fields:
                    f.set_instance_transforms(slots, worlds)
            print(f"case: tick graph, precision {precision}, depth {depth_oct}, stage timing {timing}, update {step}",
                  flush=True)
            for f in fields:
                f.tick(**kw)
            s = fields[0].stats()
            replays += int(s.tick_replayed)
            if s.tick_replayed and timing:
                assert s.stage_timed == 1
                assert s.ms_walk_primary > 0 and s.ms_walk_shadow > 0 and s.ms_blend > 0 and s.ms_total > 0
            elif s.tick_replayed:
                assert s.stage_timed == 0 and s.ms_total == 0 and s.ms_walk_shadow == 0
            e = fields[1].stats()
            assert e.tick_replayed == 0 and e.stage_timed == 1 and e.ms_total > 0
            a, b = _state(fields[0]), _state(fields[1])
            for x, y in zip(a, b):
                assert np.array_equal(x, y), f"update {step}"
            # eager work on the graph field's stream between replays
            ro = rng.uniform(0, 20, (500, 3)).astype(np.float32)
            rd = rng.normal(size=(500, 3)).astype(np.float32)
            rd /= np.linalg.norm(rd, axis=1, keepdims=True)
            q0, q1 = fields[0].query(ro, rd), fields[1].query(ro, rd)
            for x, y in zip(q0, q1):
                assert np.array_equal(x, y)
        # updates 0 (needs_clear), 1 and 5 (new arguments) capture; the others replay
        assert replays == 6
    finally:
        for f in fields:
            f.close()
