"""Receiver sampling on the GPU (sbgi_sample_irradiance, run on an MI355X with -m gpu).

The kernel samples the field's own RGBA16F atlases; the oracle's FP32 restatement samples
the same atlases read back, so the comparison isolates the gather. Bar: within 1e-5
relative of the FP32 restatement (max, with a floor of 1e-3 of the mean for near-black
texels) and 1e-4 of the FP64 oracle (p99.9), with the leak controls off and on, one and
two cascades, 8x8 and 16x16 depth tiles.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from speedball_amd import _native as N
from speedball_amd import scenegen as sg
from speedball_amd.field import ProbeField
from tests.parity import rel_err, summary

pytestmark = pytest.mark.gpu


def field_for(w, cascades=None, depth_oct=6, updates=3):
    casc = cascades or [(w.grid_min, w.grid_size, w.res)]
    f = ProbeField(casc[0][2], casc[0][0], casc[0][1], rays=w.rays, depth_oct=depth_oct, cascades=casc,
                   traversal=N.TRAV_MT_ORDERED)
    f.set_scene(w.scene)
    for _ in range(updates):
        for c in range(len(casc)):
            f.tick(cascade=c, frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
    return f


def field_state(f):
    n = len(f.cascades)
    return ([f.read_atlas(cascade=c) for c in range(n)], [f.read_atlas(depth=True, cascade=c) for c in range(n)],
            [f.read(N.BUF_META, np.float32, c) for c in range(n)])


def points(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(lo, hi, (n, 3)).astype(np.float32), rng.normal(size=(n, 3)).astype(np.float32), \
        rng.normal(size=(n, 3)).astype(np.float32)


def check(f, depth_oct, pos, nv, nd, **params):
    atl, dat, meta = field_state(f)
    assert all(a[..., :3].max() > 0 for a in atl), "the field is black"
    got, ms = f.sample(pos, nv, nd, **params)
    p = N.default_sample_params(**params)
    k = params.get("cascades", 0) or len(f.cascades)
    casc = [(gm, gs, rs) for gm, gs, rs in f.cascades][:k]
    atl, dat, meta = atl[:k], dat[:k], meta[:k]
    r32 = oracle.sample_irradiance(casc, depth_oct, atl, dat, meta, p, pos, nv, nd, prec="32")
    r64 = oracle.sample_irradiance(casc, depth_oct, atl, dat, meta, p, pos.astype(np.float64), nv.astype(np.float64),
                                   None if nd is None else nd.astype(np.float64), prec="64")
    e32 = rel_err(got, r32, floor_frac=1e-3)
    e64 = rel_err(got, r64, floor_frac=1e-3)
    print(f"\n{params}: {len(pos)} points in {ms:.3f} ms; vs FP32 {summary(e32)} vs FP64 {summary(e64)}")
    assert np.all(np.isfinite(got)) and got.max() > 0
    assert e32.max() <= 1e-5
    assert np.percentile(e64, 99.9) <= 1e-4
    return got


@pytest.mark.parametrize("depth_oct", [6, 14])
def test_sample_room(depth_oct):
    w = sg.room_boxes(res=(8, 8, 8), rays=64)
    print(f"case: sample room depth_oct {depth_oct}", flush=True)
    f = field_for(w, depth_oct=depth_oct)
    try:
        pos, nv, nd = points(20000, -0.5, 10.5, 1)
        check(f, depth_oct, pos, nv, None)
        check(f, depth_oct, pos, nv, nd, cheby_strength=1.0, cheby_bias=(0.05, 0, 0, 0))
        # probe states and relocation offsets (the classify pass's output) weight and move
        # the taps when classifyStrength is on
        rng = np.random.default_rng(7)
        meta = np.zeros((w.probes, 4), np.float32)
        meta[:, 0] = rng.uniform(0.2, 1.0, w.probes)
        meta[:, 1:] = rng.uniform(-0.2, 0.2, (w.probes, 3))
        f.write(N.BUF_META, meta.reshape(-1))
        check(f, depth_oct, pos, nv, nd, cheby_strength=0.5, classify_strength=1.0, normal_bias=0.1)
    finally:
        f.close()


def test_sample_two_cascades():
    w = sg.room_boxes(res=(6, 6, 6), rays=64)
    casc = [(w.grid_min, w.grid_size, (6, 6, 6)), ((2.0, 2.0, 2.0), (5.0, 5.0, 5.0), (6, 6, 6))]
    print("case: sample two cascades", flush=True)
    f = field_for(w, cascades=casc)
    try:
        pos, nv, nd = points(20000, 0.0, 10.0, 2)
        check(f, 6, pos, nv, nd, cheby_strength=1.0)
        # one cascade only, when asked
        check(f, 6, pos[:2000], nv[:2000], nd[:2000], cascades=1)
    finally:
        f.close()


def test_sample_rejects_bad_requests():
    w = sg.room_boxes(res=(4, 4, 4), rays=32)
    f = field_for(w, updates=1)
    try:
        pos, nv, _ = points(10, 0, 10, 3)
        with pytest.raises(RuntimeError):
            f.sample(pos, nv, cascades=2)
        out, ms = f.sample(pos[:0], nv[:0])
        assert out.shape == (0, 3)
    finally:
        f.close()
