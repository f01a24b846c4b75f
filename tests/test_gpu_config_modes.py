"""GPU parity in the configurations' own modes (run on an MI355X with -m gpu).

The instanced field at full size (32x16x24 probes x 256 rays, 261,514 triangles, 16 lights)
with the 16x16 depth tile (14x14 interior), the mode the headline update runs in, compared
with the oracle over both chained updates on a stratified probe sample:
  * trace: hit ids against the FP32 restatement (identical except near ties and edge
    grazes, counted), hit t within 1e-5, radiance p99.9 within 1e-4 (update 2 reads the
    GPU's own update-1 atlas, handed to the oracle);
  * blend: the FP64 oracle blend of the SAME GPU rays and the GPU's own previous state,
    irradiance and depth moments, for the FP32, BF16x3, FP16 and BF16 blends (the MFMA
    kernels' R = 256 K loop and the 14x14 depth contraction);
  * upload: the oracle's filter and gutters over the GPU's state, both atlases, one FP16
    step.
Also: an MFMA blend at R = 48 (a K block padded with zero weights), the SH-9 sky's l = 1, 2
bands on misses, and sbgi_poll on a queue longer than its event ring.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from speedball_amd import _native as N
from speedball_amd import scenegen as sg
from speedball_amd.field import ProbeField
from tests.parity import classify_mismatches, interior_irr, rel_err, summary

pytestmark = pytest.mark.gpu

# three probe ranges across the field: a bottom corner row, the middle, the top far rows
SAMPLE = ((0, 160), (6100, 160), (12100, 160))
PRECISIONS = {"fp32": (N.BLEND_FP32, 1e-4, 1e-3), "bf16x3": (N.BLEND_BF16X3, 1e-3, 3e-3),
              "fp16": (N.BLEND_FP16, 1e-3, 3e-3), "bf16": (N.BLEND_BF16, 1e-2, 3e-2)}


@pytest.fixture(scope="module")
def config2():
    return sg.instanced()


def depth_interior(d, n=14):
    tile = n + 2
    idx = [(ly + 1) * tile + lx + 1 for ly in range(n) for lx in range(n)]
    return d[:, idx, :]


@pytest.fixture(scope="module", params=sorted(PRECISIONS))
def config2_run(config2, request):
    """Two full-field updates on the GPU; per update the rays, hits, state and atlases."""
    w = config2
    prec = PRECISIONS[request.param][0]
    print(f"case: config-2 field {w.res} x {w.rays}, depth 14, blend {request.param}", flush=True)
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=14, precision=prec, keep_hit_ids=True,
                   traversal=N.TRAV_MT_ORDERED)
    f.set_scene(w.scene)
    steps = []
    try:
        for _ in range(2):
            f.tick(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
            steps.append(dict(rays=f.read_rays(), hits=f.read_hit_ids(), irr=f.read_irr(), depth=f.read_depth(),
                              atlas=f.read_atlas(), datlas=f.read_atlas(depth=True)))
    finally:
        f.close()
    return request.param, steps


def _oracle(w, prec):
    return oracle.OracleField(w.scene, w.grid_min, w.grid_size, w.res, w.rays, 14, prec=prec)


def test_config2_trace_vs_fp32(config2, config2_run):
    """Hit ids, t and radiance of both updates on the sample, against the FP32 restatement;
    the second update bounces off the GPU's own first atlas in both."""
    w = config2
    name, steps = config2_run
    if name != "fp32":
        pytest.skip("the trace does not depend on the blend precision")
    o = _oracle(w, "32")
    n_rays = n_mism = n_flip = 0
    for step in range(2):
        if step == 1:
            o.atlas[:] = steps[0]["atlas"].reshape(-1)
        for off, n in SAMPLE:
            p = o.params(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
            p.probe_offset, p.updated_count = off, n
            o.trace(p, n)
            g = steps[step]
            gh, gr = g["hits"][off:off + n], g["rays"][off:off + n]
            rh, rr = o.hits_view(), o.rays_view()
            mism = np.any(gh != rh, axis=-1)
            if mism.any():
                slots, ks = np.nonzero(mism)
                ties, grazes = classify_mismatches(oracle, w.scene, w, slots + off, ks, gh[slots, ks], rh[slots, ks],
                                                   gr[slots, ks, 3], edge=0.0)
                assert int((~ties).sum()) == 0  # only near ties may differ from the FP32 restatement
            same = ~mism
            hit = same & (rh[..., 0] >= 0)
            assert rel_err(gr[..., 3][hit], rr[..., 3][hit], 0).max() <= 1e-5
            e = rel_err(gr[..., :3][same], rr[..., :3][same])
            assert np.percentile(e, 99.9) <= 1e-4
            n_rays += mism.size
            n_mism += int(mism.sum())
            n_flip += int((e > 1e-3).sum())
    print(f"\nconfig 2 trace: {n_rays} rays, id mismatches {n_mism}, radiance outliers {n_flip}")
    assert n_mism <= max(8, n_rays // 20000)
    assert n_flip <= n_rays // 10000


def test_config2_blend_vs_fp64(config2, config2_run):
    """Irradiance and 14x14 depth moments of both updates: the GPU blend against the FP64
    oracle blend of the same GPU rays and the same previous state."""
    w = config2
    name, steps = config2_run
    _, bar99, barmax = PRECISIONS[name]
    o = _oracle(w, "64")
    # the blend's weights take the reference's f32 ray directions (rayDir evaluates phi in
    # f32, js/gi_probes.js:1702, 2331), so the comparison isolates the blend's arithmetic
    oracle.set_f32_dirs(True)
    try:
        _blend_steps(w, name, steps, o, bar99, barmax)
    finally:
        oracle.set_f32_dirs(False)


def _blend_steps(w, name, steps, o, bar99, barmax):
    for step in range(2):
        if step == 1:
            o.irr[:] = steps[0]["irr"].reshape(-1)
            o.depth[:] = steps[0]["depth"].reshape(-1)
        else:
            o.irr[:] = 0.0
            o.depth[:] = 0.0
        ei, ed = [], []
        for off, n in SAMPLE:
            p = o.params(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
            p.probe_offset, p.updated_count = off, n
            o.ray_data = steps[step]["rays"][off:off + n].astype(np.float64).reshape(-1).copy()
            o.blend(p)
            gi = interior_irr(steps[step]["irr"][off:off + n])[..., :3]
            ri = interior_irr(o.irr_view()[off:off + n])[..., :3]
            ei.append(rel_err(gi, ri).reshape(-1))
            gd = depth_interior(steps[step]["depth"][off:off + n])
            rd = depth_interior(o.depth_view()[off:off + n])
            ed.append(rel_err(gd, rd).reshape(-1))
        ei, ed = np.concatenate(ei), np.concatenate(ed)
        print(f"\n{name} update {step + 1}: irradiance vs FP64 {summary(ei)}; depth moments {summary(ed)}")
        assert np.percentile(ei, 99.9) <= bar99 and ei.max() <= barmax
        assert np.percentile(ed, 99.9) <= bar99 and ed.max() <= barmax


def test_config2_atlases(config2, config2_run):
    """The upload of both updates: the oracle's filter and mirrored gutters over the GPU's
    own state give the GPU's irradiance and 16x16 depth atlas tiles within one FP16 step."""
    w = config2
    name, steps = config2_run
    if name not in ("fp32", "bf16x3"):
        pytest.skip("the upload does not depend on the blend precision")
    o = _oracle(w, "64")
    rx, ry = w.res[0], w.res[1]
    for step in range(2):
        o.irr[:] = steps[step]["irr"].reshape(-1)
        o.depth[:] = steps[step]["depth"].reshape(-1)
        for off, n in SAMPLE:
            p = o.params(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
            p.probe_offset, p.updated_count = off, n
            o.upload(p)
            for probe in range(off, off + n, 7):
                col, row = probe % rx, (probe // (rx * ry)) * ry + (probe // rx) % ry
                for atlas, tile, key in ((o.atlas_view(), 8, "atlas"), (o.atlas_view(depth=True), 16, "datlas")):
                    r = atlas[row * tile:(row + 1) * tile, col * tile:(col + 1) * tile]
                    g = steps[step][key][row * tile:(row + 1) * tile, col * tile:(col + 1) * tile]
                    step16 = np.abs(np.spacing(np.float16(r))).astype(np.float64)
                    assert np.all(np.abs(g - r) <= step16 + 1e-30), (step, probe, key)


@pytest.mark.parametrize("precision", [N.BLEND_BF16X3, N.BLEND_FP16])
def test_mfma_blend_rays_48(precision):
    """R = 48 leaves the second 32-ray K block half empty: its weights must be zero, so the
    MFMA blend equals the FP64 oracle blend of the same rays within the room bars."""
    w = sg.room_boxes(res=(8, 8, 8), rays=48)
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=48, depth_oct=14, precision=precision)
    f.set_scene(w.scene)
    print(f"case: rays 48, precision {precision}", flush=True)
    try:
        f.tick(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
        rays, irr, depth = f.read_rays(), f.read_irr(), f.read_depth()
    finally:
        f.close()
    o = oracle.OracleField(w.scene, w.grid_min, w.grid_size, w.res, 48, 14, prec="64")
    o.ray_data = rays.astype(np.float64).reshape(-1).copy()
    o.blend(o.params(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis))
    e = rel_err(interior_irr(irr)[..., :3], interior_irr(o.irr_view())[..., :3])
    ed = rel_err(depth_interior(depth), depth_interior(o.depth_view()))
    print(f"rays 48: irradiance {summary(e)} depth {summary(ed)}")
    assert np.percentile(e, 99.9) <= 1e-3 and e.max() <= 3e-3
    assert np.percentile(ed, 99.9) <= 1e-3 and ed.max() <= 3e-3


def test_sky_sh9_bands_on_the_gpu():
    """Probes above an open court: misses take the SH-9 sky with non-zero l = 1 and l = 2
    coefficients (js/gi_probes.js:2189-2207). The miss radiance equals the FP32
    restatement's, and the higher bands change it (against the DC term alone) on most
    misses."""
    from tests.test_gpu_parity import open_court
    w = open_court()
    rng = np.random.default_rng(5)
    sh = np.concatenate([[0.8, 0.9, 1.0], rng.uniform(-0.3, 0.3, 24)]).astype(np.float32)
    kw = dict(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis, sky_sh=sh.tolist(), sky_intensity=1.5)
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=w.depth_oct, keep_hit_ids=True,
                   traversal=N.TRAV_MT_ORDERED)
    f.set_scene(w.scene)
    print("case: SH-9 sky over the open court", flush=True)
    try:
        f.tick(**kw)
        g_rays, g_hits = f.read_rays(), f.read_hit_ids()
    finally:
        f.close()
    o = oracle.OracleField(w.scene, w.grid_min, w.grid_size, w.res, w.rays, w.depth_oct, prec="32")
    o.update(o.params(**kw))
    miss = (o.hits_view()[..., 0] < 0) & (g_hits[..., 0] < 0)
    assert miss.mean() > 0.1
    e = rel_err(g_rays[..., :3][miss], o.rays_view()[..., :3][miss])
    print("sky misses", int(miss.sum()), "radiance vs FP32", summary(e))
    assert np.percentile(e, 99.9) <= 1e-4 and e.max() <= 1e-3
    dc = np.maximum(sh[0:3] * 0.282095, 0) * 1.5
    moved = np.abs(g_rays[..., :3][miss] - dc).max(axis=-1) > 1e-2
    assert moved.mean() > 0.5


def test_poll_beyond_the_ticket_ring(config2):
    """sbgi_poll of a ticket older than the 16-event ring answers from the oldest event in
    the ring, which completes after it: while that event is pending the old ticket reads
    as pending, never as complete before a later ticket; once the queue drains every
    ticket reads complete."""
    w = config2
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=14, precision=N.BLEND_BF16X3,
                   traversal=N.TRAV_MT_ORDERED)
    f.set_scene(w.scene)
    print("case: 20 queued ticks, poll", flush=True)
    try:
        tickets = [f.tick(wait=False, frame_jitter=w.frame_jitter, hysteresis=w.hysteresis) for _ in range(20)]
        assert tickets == list(range(tickets[0], tickets[0] + 20))
        t1, t5, t20 = tickets[0], tickets[4], tickets[19]
        first = (f.poll(t1), f.poll(t5), f.poll(t20))
        # 20 full-field updates take tens of ms: the queue is still running
        assert first[2] is False
        # a completed older ticket implies a completed newer one is not claimed before it
        assert not (first[0] and not first[1])
        f.wait(t20)
        assert all(f.poll(t) for t in tickets)
    finally:
        f.close()
