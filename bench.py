"""Benchmark: full-field DDGI probe update on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision bf16x3|bf16|fp16|fp32]

Workload: the instanced two-level scene — 8 procedural meshes (UV spheres, tori,
fbm-displaced grids) instanced 256 times with random TRS, 261,514 triangles (262,144 +-1%)
in a 40x20x25 m box,
16 point lights (log-uniform 1..100), 32x16x24 = 12,288 probes x 256 rays, 8x8
irradiance and 16x16 depth tiles, 2 bounces (each step is one update that reads the
previous update's atlas). A step is one full-field update: ray directions + weights,
trace (closest hit + NEE shadow rays + bounce), blend, upload — everything resident in
HBM. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "probe rays/s + ms per DDGI update at ≤1e-3 rel err vs FP64; fwd+adjoint ms"
PRECISIONS = {"fp32": 0, "bf16": 1, "bf16x3": 2, "fp16": 3}
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def walk_bytes(blas, tlas, inst, tris) -> float:
    """Reference-layout bytes a traversal touches: 32 B per BLAS node, 48 B per TLAS node,
    64 B per instance entry, 36 B per triangle test (12 B index + 3 x 8 B vertex slots)."""
    return 32.0 * blas + 48.0 * tlas + 64.0 * inst + 36.0 * tris


def algorithmic_bytes(stats, hits: int) -> float:
    """Whole trace: every walk's node/triangle bytes, plus per primary hit 16 B payload +
    112 B material + 256 B of bounce taps."""
    return walk_bytes(stats.blas_nodes, stats.tlas_nodes, stats.inst_visits, stats.tri_tests) + (16 + 112 + 256) * hits


def shadow_phases(stats) -> dict:
    """Wave steps of the shadow walk's NODE, TRI, INST phases and refills (counting run),
    and the share of the 64 lanes that did work in each: node visits, triangle tests and
    instance entries per lane-step (INST steps that return to the TLAS count as idle)."""
    w = [int(x) for x in stats.shadow_waves]
    work = (stats.shadow_blas_nodes + stats.shadow_tlas_nodes, stats.shadow_tri_tests, stats.shadow_inst_visits)
    out = {}
    for name, steps, done in zip(("node", "tri", "inst"), w[:3], work):
        out[name] = {"wave_steps": steps, "lanes_busy": float(done) / (64.0 * steps) if steps else 0.0}
    out["refill"] = {"wave_steps": w[3]}
    return out


def shadow_walk_bytes(stats) -> float:
    """The shadow any-hit walk alone (the dominant kernel)."""
    return walk_bytes(stats.shadow_blas_nodes, stats.shadow_tlas_nodes, stats.shadow_inst_visits,
                      stats.shadow_tri_tests)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "walk_instanced_pmc.md")
# shadow rays walk the light-space candidate lists (False: --no-shadow-lists, the TLAS)
SHADOW_LISTS = True


def measured_traffic(kernel_prefix: str):
    """HBM bytes per dispatch (calibrated FETCH_SIZE + WRITE_SIZE) of a kernel, from the
    committed rocprofv3 --pmc summary of this workload; None when it is not there."""
    v = _summary_lines(TRAFFIC_FILE, "traffic", kernel_prefix)
    return float(v[0][0]) + float(v[0][1]) if v else None


def cpu_baseline(w, seconds_target=12.0):
    """FP32 restatement (OpenMP over probes) on a bounded sample of the same field: the
    first n probes (n grows x4 until the run takes a few seconds), then whole-field
    updates repeated until about `seconds_target` seconds of CPU work."""
    import oracle
    o = oracle.OracleField(w.scene, w.grid_min, w.grid_size, w.res, w.rays, w.depth_oct, prec="32")
    p = o.params(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)

    def one(n):
        p.probe_offset = 0
        p.updated_count = n
        t0 = time.perf_counter()
        o.trace(p, n)
        o.blend(p)
        o.upload(p)
        return time.perf_counter() - t0

    n = 16
    dt = one(n)
    while dt * 4 < seconds_target and n < w.probes:
        n = min(n * 4, w.probes)
        dt = one(n)
    reps, total = 1, dt
    while total < seconds_target and n == w.probes and reps < 16:
        total += one(n)
        reps += 1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    return {"value": reps * n * w.rays / total, "unit": "rays/s", "cores": threads, "kind": "port",
            "sample": f"FP32 restatement (oracle/), OpenMP on {threads} threads: {reps} update(s) of the first {n} "
                      f"probes x {w.rays} rays of the same field from cleared state, {total:.2f} s"}


def sample_ranges(total: int, n_ranges: int = 8, per: int = 48):
    """Stratified probe sample: n_ranges runs of `per` consecutive probes spread evenly over
    the field (each run covers one stretch of x rows at its own y and z)."""
    return [(int(i * total / n_ranges), per) for i in range(n_ranges)]


def fp64_check(w, states, depth_oct, ranges):
    """Irradiance of both GPU updates against the FP64 oracle on a stratified probe sample.

    Update 1 starts from cleared state; update 2 is recomputed by the oracle from the GPU's
    own update-1 atlas and temporal state (states[0]), so each update is checked on its
    own. Each update is compared twice: with the oracle's exact ray directions, and with the
    reference's f32 directions (phi = k GA rounded in f32, js/gi_probes.js:1702) widened to
    double. The second isolates the arithmetic of the GPU path; the difference between the
    two is the f32 direction definition's share. Texels off by more than 1e-3 are counted,
    and those that the f32 directions bring within 1e-3 are counted as explained."""
    import oracle
    from tests.parity import rel_err, interior_irr  # shared metric definition
    out = []
    for u in range(2):
        errs = {}
        for mode in ("exact", "f32"):
            oracle.set_f32_dirs(mode == "f32")
            try:
                o = oracle.OracleField(w.scene, w.grid_min, w.grid_size, w.res, w.rays, depth_oct, prec="64")
                if u == 1:
                    o.atlas[:] = states[0]["atlas"].reshape(-1)
                    o.irr[:] = states[0]["irr"].reshape(-1)
                    o.depth[:] = states[0]["depth"].reshape(-1)
                # every range's bounce reads the atlas the update started from (the GPU
                # traces all probes before any upload), so each range starts from that atlas
                atlas0 = o.atlas.copy()
                e = []
                for off, n in ranges:
                    o.atlas[:] = atlas0
                    p = o.params(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
                    p.probe_offset, p.updated_count = off, n
                    o.update(p)
                    g = interior_irr(states[u]["irr"][off:off + n])[..., :3]
                    r = interior_irr(o.irr_view()[off:off + n])[..., :3]
                    e.append(rel_err(g, r).max(axis=-1).reshape(-1))
                errs[mode] = np.concatenate(e)
            finally:
                oracle.set_f32_dirs(False)
        ex, f3 = errs["exact"], errs["f32"]
        over = ex > 1e-3
        out.append({"update": u + 1, "p99": float(np.percentile(ex, 99)), "max": float(ex.max()),
                    "p99_f32_dirs": float(np.percentile(f3, 99)), "max_f32_dirs": float(f3.max()),
                    "texels": int(ex.size), "texels_over_1e-3": int(over.sum()),
                    "explained_by_f32_dirs": int((over & (f3 <= 1e-3)).sum())})
    return {"probes_checked": int(sum(n for _, n in ranges)), "sample": f"{len(ranges)} runs of "
            f"{ranges[0][1]} consecutive probes spread over the field", "updates": out}


def read_state(f):
    return {"irr": f.read_irr(), "depth": f.read_depth(), "atlas": f.read_atlas()}


def room_leg(steps=20, warmup=3):
    """The room workload: room + 64 boxes (780 triangles), a point
    and a directional light, 16^3 probes x 64 rays, plain FP32 blend, in the configuration
    depth mode (16x16 tile) and the reference-parity mode (depth in the 8x8 tile). Per mode:
    ms per update (wall clock over `steps` queued updates), the stage split, and the FP64
    check of both updates on the whole field."""
    from speedball_amd import scenegen
    from speedball_amd.field import ProbeField
    import torch
    out = {"workload": "room: 10 m room + 64 boxes (780 triangles), 1 point + 1 directional light, 16^3 probes x "
                       "64 rays, FP32 blend"}
    for depth_oct in (14, 6):
        w = scenegen.room_boxes(depth_oct=depth_oct)
        f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=depth_oct, precision=0, traversal=1,
                       shadow_lists=SHADOW_LISTS)
        f.set_scene(w.scene)
        kw = dict(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
        states = []
        for _ in range(2):
            f.tick(**kw)
            states.append(read_state(f))
        for _ in range(warmup):
            f.tick(**kw)
        f.wait()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            f.tick(wait=False, **kw)
        f.wait()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / steps
        f.set_stage_timing(True)  # stage events for one more tick (they cost ~0.1 ms a tick)
        f.tick(**kw)
        s = f.stats()
        f.close()
        chk = fp64_check(w, states, depth_oct, [(0, w.probes)])
        out["depth_16x16" if depth_oct == 14 else "depth_8x8_reference"] = {
            "ms_per_update": ms, "probe_rays_per_s": w.probes * w.rays / (ms / 1000.0),
            "stage_ms": {"trace": s.ms_trace, "blend": s.ms_blend, "upload": s.ms_upload, "total": s.ms_total},
            "irradiance_rel_err_vs_fp64": chk}
    return out


# ---------------------------------------------------------------- cache and issue ceilings
VALU_CLASSES = ("add_f32", "mul_f32", "fma_f32", "trans_f32", "int32", "cvt", "int64")
CEILINGS_FILE = os.path.join(ROOT, "profiles", "roofline_ceilings.md")


def _summary_lines(path: str, tag: str, kernel_prefix: str = None):
    """Fields of the machine-readable lines `tag [kernel] values...` of a committed summary."""
    out = []
    try:
        with open(path) as fh:
            for line in fh:
                parts = line.split()
                if not parts or parts[0] != tag:
                    continue
                if kernel_prefix is None:
                    out.append(parts[1:])
                elif len(parts) > 2 and parts[1].startswith(kernel_prefix):
                    out.append(parts[2:])
    except OSError:
        return []
    return out


def ceilings():
    """Measured ceilings of this GPU (speedball_amd/calib/ceilings, summarised in
    profiles/roofline_ceilings.md): VALU class rates and peak (wave-instr/s), the L2 read
    ceiling (bytes/s of 128-byte requests) and the L1 ceiling (tag accesses/s)."""
    out = {}
    for f in _summary_lines(CEILINGS_FILE, "rate"):
        out["rate_" + f[0]] = float(f[1]) * 1e9
    for tag in ("l2_ceiling", "l1_ceiling"):
        v = _summary_lines(CEILINGS_FILE, tag)
        if v:
            out[tag] = float(v[0][0]) * 1e9
    return out


def valu_issue(kernel_prefix: str, launch_ms: float, ceil: dict):
    """VALU issue of a kernel from the committed counter summary: wave-instructions per
    launch (SQ_INSTS_VALU) over the launch time, against the fastest measured issue rate
    (a lower bound on the share of the issue capacity used, whatever the mix), and the
    bracket of the additive class model (each counted class at its own measured rate, the
    unclassified ones at the fast or the slow rate)."""
    mix = _summary_lines(TRAFFIC_FILE, "valu_mix", kernel_prefix)
    if not mix or "rate_peak" not in ceil:
        return None
    v = [float(x) for x in mix[0]]
    total, by = v[0], dict(zip(VALU_CLASSES + ("other",), v[1:]))
    s = launch_ms / 1000.0
    base = sum(by[c] / ceil["rate_" + c] for c in VALU_CLASSES)
    lo = (base + by["other"] / ceil["rate_other_fast"]) / s
    hi = (base + by["other"] / ceil["rate_other_slow"]) / s
    return {"achieved": total / s / 1e9, "peak": ceil["rate_peak"] / 1e9, "unit": "G wave-instr/s",
            "frac": total / s / ceil["rate_peak"], "mix_model_frac": [lo, hi],
            "instructions_per_launch": total, "mix": by}


def l2_requests(kernel_prefix: str, launch_ms: float, ceil: dict):
    """L2->L1 request bytes per launch (TCP_TCC_READ_REQ x 128 B) over the L2 read ceiling."""
    v = _summary_lines(TRAFFIC_FILE, "l2_req", kernel_prefix)
    if not v or "l2_ceiling" not in ceil:
        return None
    b = float(v[0][0])
    gbs = b / (launch_ms / 1000.0) / 1e9
    return {"bytes_per_launch": b, "gbs": gbs, "frac": gbs * 1e9 / ceil["l2_ceiling"]}


def mfma_util(kernel_prefix: str):
    """MFMA busy fraction of a kernel from the committed counter summary:
    SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE per XCD x 1,024 SIMDs)."""
    v = _summary_lines(TRAFFIC_FILE, "mfma_busy", kernel_prefix)
    return float(v[0][0]) if v else None


def adjoint_leg(reps=5):
    """Forward + hand-written adjoint of the differentiable two-update solve on its own
    workload (4 tori x 1024 triangles x 64 instances, 16 lights, 16^3 probes x 128 rays):
    device time from the library's HIP events, median of `reps`."""
    from speedball_amd import scenegen
    from speedball_amd.field import ProbeField
    w, albedo, target = scenegen.differentiable()
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=6, precision=0, traversal=1,
                   shadow_lists=SHADOW_LISTS)
    f.set_scene(w.scene)
    params = dict(frame_jitter=w.frame_jitter, hysteresis=0.0)
    g = 2 * (f.diff_forward(albedo, **params) - f.diff_forward(target, **params))
    fwd, adj = [], []
    for _ in range(reps):
        f.diff_forward(albedo, **params)
        f.diff_adjoint(g)
        a, b = f.diff_timing()
        fwd.append(a)
        adj.append(b)
    n_alb, n_l = f.diff_param_count()
    f.close()
    fm, am = float(np.median(fwd)), float(np.median(adj))
    return {"workload": "differentiable: 4 tori x 1024 triangles in 64 instances (65536 albedo entries), 16 point "
                        "lights, 16^3 probes x 128 rays, 2 updates", "albedo_params": n_alb, "light_params": n_l,
            "forward_ms": fm, "adjoint_ms": am, "fwd_plus_adjoint_ms": fm + am, "ratio_to_forward": (fm + am) / fm}


def dynamic_leg(frames=600, warmup=10):
    """Dynamic scene with per-frame device refit: 32 skinned tubes (4-bone LBS, 1,048,576
    triangles) + 128 rigid instances on a random walk; every frame: skinning + BLAS refit,
    instance records + bounds + SAH TLAS rebuild, then a full update of 4 cascades of 24^3
    probes x 128 rays (BF16x3 blend, 2 bounces). Device time from the library's HIP events;
    the irradiance error is the first update of the largest cascade on the final frame's
    geometry against the FP64 oracle (24 probes)."""
    import math
    import oracle
    from speedball_amd import _native as N
    from speedball_amd import scenegen
    from speedball_amd.field import ProbeField
    from tests.parity import rel_err, interior_irr
    t0 = time.perf_counter()
    w = scenegen.dynamic()
    f = ProbeField(w.cascades[0][2], w.cascades[0][0], w.cascades[0][1], rays=w.rays, depth_oct=w.depth_oct,
                   precision=PRECISIONS["bf16x3"], cascades=w.cascades, traversal=1, shadow_lists=SHADOW_LISTS)
    f.set_stage_timing(True)  # the frames are timed by the field's stage events
    f.set_scene(w.scene)
    f.set_instance_transforms(np.arange(w.scene.inst_count), list(w.slot_worlds))
    f.set_skins(w.skin_sets, w.n_bones)
    log(f"dynamic: scene {w.scene.tri_count} triangles, setup {time.perf_counter() - t0:.1f} s")
    # the whole animation is generated up front so the loop only enqueues device work
    total = warmup + frames
    bones = [w.bones(i / w.fps) for i in range(total)]
    rigid = [w.rigid_step() for _ in range(total)]
    params = dict(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)
    rec = {"skin": [], "refit": [], "tlas": [], "update": [], "wall": []}
    for i in range(total):
        ts = time.perf_counter()
        f.skin(bones[i], update_tlas=False)
        f.set_instance_transforms(*rigid[i])
        upd = 0.0
        for c in range(len(w.cascades)):
            f.tick(cascade=c, **params)
            upd += f.stats().ms_total
        sk, rf, tl = f.edit_timing()
        wall = (time.perf_counter() - ts) * 1e3
        if i >= warmup:
            for k, v in zip(("skin", "refit", "tlas", "update", "wall"), (sk, rf, tl, upd, wall)):
                rec[k].append(v)
        if (i + 1) % 100 == 0:
            log(f"dynamic: frame {i + 1}/{total}, update {upd:.2f} ms, edits {sk + rf + tl:.3f} ms")
    # device rebuild of every skinned BLAS on the final pose (the refit keeps the rest
    # pose's tree): its time, and the update time before / after it
    def update_ms():
        tot = 0.0
        for c in range(len(w.cascades)):
            f.tick(cascade=c, **params)
            tot += f.stats().ms_total
        return tot

    before = float(np.median([update_ms() for _ in range(3)]))
    tube_ranges = [(int(br), int(be)) for (_, _, br, be, _, _) in w.skin_sets]
    f.rebuild_blas(tube_ranges)  # warm-up (first-use allocations)
    ts = time.perf_counter()
    f.rebuild_blas(tube_ranges)
    rebuild_wall = (time.perf_counter() - ts) * 1e3
    _, rebuild_dev, rebuild_tlas = f.edit_timing()
    after = float(np.median([update_ms() for _ in range(3)]))
    rebuild = {"blases": len(tube_ranges), "triangles": int(w.scene.meta["skinned_triangles"]),
               "device_ms": rebuild_dev + rebuild_tlas, "wall_ms": rebuild_wall,
               "update_ms_refit_tree": before, "update_ms_rebuilt_tree": after}
    # error statement: first update (cleared state, empty bounce atlas) of the largest
    # cascade on the final frame's geometry, in a fresh field
    big = len(w.cascades) - 1
    s = w.scene
    vd = f.read(N.BUF_VERTEX_DATA, np.float32)
    nodes = f.read(N.BUF_BVH_NODES, np.uint32)
    mats = f.read(N.BUF_MATERIALS, np.float32).copy()
    b = f.read(N.BUF_INST_BOUNDS, np.float32)
    tl = mats[s.tlas_base:s.tlas_base + s.tlas_node_count * 12].copy()
    assert oracle.refit_tlas(tl, b.copy())
    mats[s.tlas_base:s.tlas_base + s.tlas_node_count * 12] = tl
    import dataclasses
    kw = {fl.name: getattr(s, fl.name) for fl in dataclasses.fields(s) if fl.name != "_view"}
    kw.update(vertex_data=vd, bvh_nodes=nodes, materials=mats, tri_index=f.read(N.BUF_TRI_INDEX, np.uint32),
              tri_material=f.read(N.BUF_TRI_MATERIAL, np.uint32))
    s2 = scenegen.Scene(**kw)
    f.close()
    gm, gs, res = w.cascades[big]
    f = ProbeField(res, gm, gs, rays=w.rays, depth_oct=w.depth_oct, precision=PRECISIONS["bf16x3"], traversal=1,
                   shadow_lists=SHADOW_LISTS)
    f.set_scene(s2)
    f.tick(**params)
    o = oracle.OracleField(s2, gm, gs, res, w.rays, w.depth_oct, prec="64")
    p = o.params(**params)
    p.updated_count = 24
    o.update(p)
    e = rel_err(interior_irr(f.read_irr())[:24, :, :3], interior_irr(o.irr_view())[:24, :, :3])
    f.close()

    def q(v):
        v = np.asarray(v)
        return {"median": float(np.median(v)), "p99": float(np.percentile(v, 99)), "max": float(v.max())}

    edits = np.array(rec["skin"]) + np.array(rec["refit"]) + np.array(rec["tlas"])
    frame = edits + np.array(rec["update"])
    rays = w.probes * w.rays
    return {"workload": "dynamic: 32 skinned tubes (4-bone LBS, +-30 deg at 0.5 Hz, 1,048,576 triangles) + 128 "
                        "rigid instances on a random walk, 16 point lights; per frame skinning + BLAS refit, instance "
                        "bounds + SAH TLAS rebuild, update of 4 cascades of 24^3 probes x 128 rays, bf16x3 blend",
            "frames": frames, "triangles": int(s.tri_count), "probes": int(w.probes), "rays_per_frame": int(rays),
            "ms_per_frame": q(frame), "refit_ms": q(edits), "update_ms": q(rec["update"]),
            "skin_ms": q(rec["skin"]), "blas_refit_ms": q(rec["refit"]), "tlas_ms": q(rec["tlas"]),
            "wall_ms_per_frame": q(rec["wall"]), "probe_rays_per_s": rays / (float(np.median(frame)) / 1e3),
            "rebuild": rebuild,
            "irradiance_rel_err_vs_fp64": {"p99": float(np.percentile(e, 99)), "max": float(e.max()),
                                           "probes_checked": 24, "cascade": big}}


def batch_leg(frames=10, autotune=True):
    """Batched uneven fields: 48 seeded room-and-boxes scenes (triangles log-uniform
    [2k, 2M], probes per cascade log-uniform [512, 32768], rays from {32, 64, 128, 256},
    1-4 cascades, bf16x3 blend, depth 16x16), every cascade of every field updated per frame
    by one batch frame (one persistent closest-hit walk + one persistent shadow walk over
    all segments, replayed from a hipGraph). Beside it, the same fields updated one cascade
    at a time with sbgi_tick (device time summed)."""
    from concurrent.futures import ThreadPoolExecutor
    import oracle
    from speedball_amd import scenegen
    from speedball_amd.field import Batch, ProbeField
    from tests.parity import rel_err, interior_irr
    t0 = time.perf_counter()
    specs = scenegen.batch_specs()
    with ThreadPoolExecutor(8) as ex:
        built = list(ex.map(lambda s: scenegen.room_boxes_scaled(depth_oct=14, **s), specs))

    def make(i):
        scene, casc, _ = built[i]
        f = ProbeField(casc[0][2], casc[0][0], casc[0][1], rays=specs[i]["rays"], depth_oct=14,
                       precision=PRECISIONS["bf16x3"], cascades=casc, traversal=1, shadow_lists=SHADOW_LISTS)
        f.set_scene(scene)
        return f

    with ThreadPoolExecutor(8) as ex:
        fields = list(ex.map(make, range(len(specs))))
    tris = sum(b[0].tri_count for b in built)
    rays = sum(f.probes(c) * f.rays for f in fields for c in range(len(f.cascades)))
    segs = sum(len(f.cascades) for f in fields)
    log(f"batch: {len(fields)} fields, {segs} segments, {tris} triangles, {rays} probe rays/frame, "
        f"setup {time.perf_counter() - t0:.1f} s")
    params = [dict(frame_jitter=b[2], hysteresis=0.97) for b in built]
    batch = Batch(fields)
    batch.tick(params)  # cleared first frame (eager)
    tuned = None
    if autotune:
        tuning, block, tms = batch.autotune(params)
        tuned = {"walk_tuning": tuning, "walk_block": block, "walk_waves_per_cu": batch.walk_waves, "frame_ms": tms}
        log(f"batch: autotuned {tuning} block {block} ({tms:.2f} ms eager)")
    batch.tick(params)  # capture
    ms, wall, rep = [], [], []
    for _ in range(frames):
        ts = time.perf_counter()
        batch.tick(params)
        wall.append((time.perf_counter() - ts) * 1e3)
        m, r = batch.timing()
        ms.append(m)
        rep.append(r)
    batch.close()
    # the same fields, one cascade at a time (timed by their stage events)
    for f in fields:
        f.set_stage_timing(True)
    single = []
    for _ in range(3):
        tot = 0.0
        for f, kw in zip(fields, params):
            for c in range(len(f.cascades)):
                f.tick(cascade=c, **kw)
                tot += f.stats().ms_total
        single.append(tot)
    for f in fields:
        f.close()
    # error statement: the first update of the smallest field's largest cascade vs FP64
    i = min(range(len(specs)), key=lambda k: built[k][0].tri_count)
    scene, casc, jit = built[i]
    gm, gs, res = casc[-1]
    f = ProbeField(res, gm, gs, rays=specs[i]["rays"], depth_oct=14, precision=PRECISIONS["bf16x3"], traversal=1,
                   shadow_lists=SHADOW_LISTS)
    f.set_scene(scene)
    f.tick(frame_jitter=jit, hysteresis=0.97)
    o = oracle.OracleField(scene, gm, gs, res, specs[i]["rays"], 14, prec="64")
    p = o.params(frame_jitter=jit, hysteresis=0.97)
    p.updated_count = 24
    o.update(p)
    e = rel_err(interior_irr(f.read_irr())[:24, :, :3], interior_irr(o.irr_view())[:24, :, :3])
    f.close()
    med = float(np.median(ms))
    return {"workload": "batched: 48 room-and-boxes scenes, triangles log-uniform [2k, 2M], probes per cascade "
                        "log-uniform [512, 32768], rays {32,64,128,256}, 1-4 cascades, bf16x3 blend, depth 16x16",
            "fields": len(fields), "segments": segs, "triangles": int(tris), "probe_rays_per_frame": int(rays),
            "frames": frames, "ms_per_frame": med, "ms_per_frame_p99": float(np.percentile(ms, 99)),
            "wall_ms_per_frame": float(np.median(wall)), "graph_replayed": bool(all(rep)),
            "probe_rays_per_s": rays / (med / 1e3), "unbatched_ms_per_frame": float(np.median(single)),
            "speedup_vs_unbatched": float(np.median(single)) / med, "autotune": tuned,
            "irradiance_rel_err_vs_fp64": {"p99": float(np.percentile(e, 99)), "max": float(e.max()),
                                           "probes_checked": 24, "field": i}}


def sample_leg(f, w, n=1 << 20, reps=5):
    """Receiver sampling (sbgi_sample_irradiance) of the updated headline field at 2^20
    random points with random normals, leak controls on (chebyStrength 1): kernel time
    from the library's events (median of reps), and the points against the FP32
    restatement over the same atlases on a 4096-point subset."""
    import oracle
    from speedball_amd import _native as N
    from tests.parity import rel_err
    rng = np.random.default_rng(11)
    pos = (rng.uniform(0, 1, (n, 3)) * np.array(w.grid_size) + np.array(w.grid_min)).astype(np.float32)
    nv = rng.normal(size=(n, 3)).astype(np.float32)
    ms = []
    out = None
    for _ in range(reps):
        out, m = f.sample(pos, nv, cheby_strength=1.0)
        ms.append(m)
    k = 4096
    ref = oracle.sample_irradiance([(w.grid_min, w.grid_size, w.res)], w.depth_oct, [f.read_atlas()],
                                   [f.read_atlas(depth=True)], [f.read(N.BUF_META, np.float32)],
                                   N.default_sample_params(cheby_strength=1.0), pos[:k], nv[:k], prec="32")
    e = rel_err(out[:k], ref, floor_frac=1e-3)
    med = float(np.median(ms))
    # per point: 8 taps x (16 B meta + 4 x 8 B depth texels + 4 x 8 B irradiance texels) + 36 B in/out
    alg = n * (8 * (16 + 32 + 32) + 36)
    return {"points": n, "ms": med, "points_per_s": n / (med / 1e3), "algorithmic_gbs": alg / (med / 1e3) / 1e9,
            "rel_err_vs_fp32": {"max": float(e.max()), "points_checked": k}}


def reflect_leg(steps=10):
    """Reflection lobes on the headline field (BF16x3 contractions): ms per update and the
    blend stage (irradiance + depth blend, plus the rough lobe and the glossy cache when on)
    for reflectionQuality off / rough / high / ultra, and the lobes' work per update."""
    from speedball_amd import scenegen
    from speedball_amd import _native as N
    from speedball_amd.field import ProbeField
    w = scenegen.instanced()
    out = {"workload": "instanced (261514 triangles, 32x16x24 probes x 256 rays), bf16x3, ordered", "steps": steps}
    for name, q in (("off", N.REFLECT_OFF), ("rough", N.REFLECT_ROUGH), ("high", N.REFLECT_HIGH),
                    ("ultra", N.REFLECT_ULTRA)):
        f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=w.depth_oct,
                       precision=PRECISIONS["bf16x3"], traversal=1, reflections=q, shadow_lists=SHADOW_LISTS)
        f.set_stage_timing(True)
        f.set_scene(w.scene)
        total, blend = [], []
        for i in range(steps + 2):
            f.tick(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis, glossy_phase=i % 2)
            if i >= 2:
                x = f.stats()
                total.append(x.ms_total)
                blend.append(x.ms_blend)
        f.close()
        G = {N.REFLECT_HIGH: 10, N.REFLECT_ULTRA: 18}.get(q, 0)
        texels = 36 * (q != N.REFLECT_OFF) + G * G // (2 if q == N.REFLECT_HIGH else 1)
        out[name] = {"ms_total": float(np.median(total)), "ms_blend_stage": float(np.median(blend)),
                     "lobe_texels_per_probe": texels,
                     "lobe_gflop": 2.0 * w.probes * w.rays * 4 * texels / 1e9}
        log(f"reflections {name}: {out[name]}")
    for name in ("rough", "high", "ultra"):
        extra = out[name]["ms_blend_stage"] - out["off"]["ms_blend_stage"]
        out[name]["lobe_ms"] = extra
        out[name]["lobe_tflops"] = out[name]["lobe_gflop"] / max(extra, 1e-6)
    return out


def lights_leg(steps=10):
    """Clustered NEE and emitters (the reference's clusteredLighting and emitter path): the
    instanced scene with 48 point lights of range 8 m and 16 emissive-sphere emitters,
    updated by a field that loops over every light and by a clustered field (per-cell
    lists over the probe grid's cells); both bf16x3, ordered traversal. Reports ms per
    update, the emission / shadow-walk / emitter-visibility / blend stages, shadow rays and
    the mean list length, plus the FP64 error of the clustered field with emitters."""
    import oracle
    from speedball_amd import scenegen
    from speedball_amd.field import ProbeField
    from tests.parity import rel_err, interior_irr
    w = scenegen.instanced(n_lights=48, light_range=8.0, n_emitters=16)
    kw = dict(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis, emitter_vis_seed=0.3)
    out = {"workload": "instanced (261514 triangles, 32x16x24 probes x 256 rays) with 48 point lights of range "
                       "8 m and 16 emitters; bf16x3 blend, ordered traversal", "steps": steps}
    first = None
    for name, clustered in (("all_lights", False), ("clustered", True)):
        f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=w.depth_oct,
                       precision=PRECISIONS["bf16x3"], traversal=1, clustered=clustered, shadow_lists=SHADOW_LISTS)
        f.set_stage_timing(True)
        f.set_scene(w.scene)
        f.tick(**kw)
        if clustered:
            first = f.read_irr()
            grid = f.read_light_grid()
            out["mean_list_length"] = float(np.mean(np.minimum(grid[:, 0], 16)))
            out["overflow_cells"] = int((grid[:, 0] > 16).sum())
        f.tick(**kw)
        st = {k: [] for k in ("total", "emit", "walk_shadow", "emitter_vis", "blend")}
        for _ in range(steps):
            f.tick(**kw)
            x = f.stats()
            st["total"].append(x.ms_total)
            st["emit"].append(x.ms_emit)
            st["walk_shadow"].append(x.ms_walk_shadow)
            st["emitter_vis"].append(x.ms_emitter_vis)
            st["blend"].append(x.ms_blend)
        f.set_counting(True)
        f.tick(**kw)
        shadow = int(f.stats().shadow_rays)
        f.set_counting(False)
        f.close()
        out[name] = dict({f"ms_{k}": float(np.median(v)) for k, v in st.items()}, shadow_rays=shadow)
        log(f"lights: {name}: {out[name]}")
    # the first update's irradiance on 24 probes against the FP32 restatement (the parity
    # bar: same f32 rays) and the FP64 oracle (these corner probes are dim, so a primary
    # ray that grazes an edge and shades the neighbouring facet moves a texel by percents)
    for prec in ("32", "64"):
        o = oracle.OracleField(w.scene, w.grid_min, w.grid_size, w.res, w.rays, w.depth_oct, prec=prec)
        o.set_clustered()
        p = o.params(**kw)
        p.updated_count = 24
        o.update(p)
        e = rel_err(interior_irr(first)[:24, :, :3], interior_irr(o.irr_view())[:24, :, :3])
        out[f"irradiance_rel_err_vs_fp{prec}"] = {"p99": float(np.percentile(e, 99)), "max": float(e.max()),
                                                  "probes_checked": 24}
    out["clustered_speedup"] = out["all_lights"]["ms_total"] / out["clustered"]["ms_total"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="bf16x3", choices=sorted(PRECISIONS))
    ap.add_argument("--traversal", default="ordered", choices=["compat", "ordered"],
                    help="compat: the reference's TLAS and left-first node order; ordered: SAH TLAS + "
                         "near-first octant node copies (same hits except exact-t ties)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 subset check")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the forward+adjoint leg")
    ap.add_argument("--no-dynamic", action="store_true", help="skip the dynamic-scene (per-frame refit) leg")
    ap.add_argument("--dynamic-frames", type=int, default=600, help="frames of the dynamic-scene leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-fields leg")
    ap.add_argument("--batch-frames", type=int, default=10, help="timed frames of the batched-fields leg")
    ap.add_argument("--no-lights", action="store_true", help="skip the clustered-lights and emitters leg")
    ap.add_argument("--no-reflect", action="store_true", help="skip the reflection-lobes leg")
    ap.add_argument("--no-room", action="store_true", help="skip the room leg (16^3 probes x 64 rays)")
    ap.add_argument("--no-shadow-lists", action="store_true",
                    help="walk the headline field's shadow rays over the TLAS instead of the light-space lists")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 256 MiB device-to-device copy kernel (reference for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (the probe field would shard "
                                   "by probe slabs with an atlas all-gather per update, not built here)"}))
        sys.exit(2)

    from speedball_amd import scenegen
    from speedball_amd.field import ProbeField
    from speedball_amd import _native as N

    global SHADOW_LISTS
    SHADOW_LISTS = not args.no_shadow_lists
    t0 = time.time()
    w = scenegen.instanced()
    log(f"scene: {w.scene.meta}, {w.scene.node_count} BLAS nodes, {w.scene.tlas_node_count} TLAS nodes, "
        f"built in {time.time() - t0:.1f}s")
    prec = PRECISIONS[args.precision]
    trav = N.TRAV_MT_ORDERED if args.traversal == "ordered" else N.TRAV_MT_COMPAT
    f = ProbeField(w.res, w.grid_min, w.grid_size, rays=w.rays, depth_oct=w.depth_oct, precision=prec, traversal=trav,
                   shadow_lists=SHADOW_LISTS)
    f.set_scene(w.scene)
    kw = dict(frame_jitter=w.frame_jitter, hysteresis=w.hysteresis)

    if args.copy_calibration:
        import torch
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # the first two updates (from cleared state, then bouncing off the first): kept for the
    # FP64 check
    states = []
    for _ in range(2):
        f.tick(**kw)
        states.append(read_state(f))
    for _ in range(max(0, args.warmup - 2)):
        f.tick(**kw)

    import torch  # device-wide synchronisation around the timed region

    # timed region: K updates queued back to back on the field's stream
    f.wait()
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        f.tick(wait=False, **kw)
    f.wait()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_step = wall * 1000.0 / args.steps

    # per-stage device time from the library's HIP events (recorded on the field's stream;
    # the timed region above runs without them: they cost ~0.1 ms a tick)
    f.set_stage_timing(True)
    stage = {"trace": [], "blend": [], "upload": [], "total": [], "raydir": [], "walk_primary": [], "emit": [],
             "walk_shadow": [], "resolve": []}
    for _ in range(args.steps):
        f.tick(**kw)
        s = f.stats()
        stage["trace"].append(s.ms_trace)
        stage["blend"].append(s.ms_blend)
        stage["upload"].append(s.ms_upload)
        stage["raydir"].append(s.ms_raydir)
        stage["total"].append(s.ms_total)
        stage["walk_primary"].append(s.ms_walk_primary)
        stage["emit"].append(s.ms_emit)
        stage["walk_shadow"].append(s.ms_walk_shadow)
        stage["resolve"].append(s.ms_resolve)
    rays = w.probes * w.rays
    value = rays / (ms_step / 1000.0)

    # traversal counters of one instrumented update (same inputs) for the algorithmic bytes
    f.set_counting(True)
    f.tick(**kw)
    cs = f.stats()
    f.set_counting(False)
    hits = int(np.count_nonzero(f.read_rays()[..., 3] >= 0))
    bytes_alg = algorithmic_bytes(cs, hits)
    trace_ms = float(np.mean(stage["trace"]))
    # roofline of the dominant kernel, the shadow any-hit walk, timed by its own events; the
    # primary walk beside it
    ceil = ceilings()
    ordered = args.traversal == "ordered"
    sw_bytes = shadow_walk_bytes(cs)
    sw_ms = float(np.mean(stage["walk_shadow"]))
    # the shadow walk over the light-space candidate lists (the field's default on this
    # cache-resident scene) or over the TLAS
    lists_on = f.nbytes(N.BUF_SHADOW_LIST) > 0 and f.walk_tuning()["shadow_lists"] == 1
    shadow_kernel = ("k_walk<1,false,true,ShadowListSrc,false,512>" if lists_on
                     else "k_walk<1,false,true,ShadowSrc,false,512>")
    pw_bytes = walk_bytes(cs.blas_nodes, cs.tlas_nodes, cs.inst_visits, cs.tri_tests)
    pw_ms = float(np.mean(stage["walk_primary"]))

    def walk_roofline(kernel, ref_bytes, ms):
        """The walk against the measured L2 read ceiling: the bytes it pulls from L2 (its
        TCP->TCC requests from the committed counter summary) over its live launch time;
        VALU issue, HBM and the reference-layout bytes beside it."""
        l2 = l2_requests(kernel, ms, ceil) if ordered else None
        l2_peak = ceil.get("l2_ceiling", 0.0) / 1e9 or None
        traffic = measured_traffic(kernel) if ordered else None
        ref_gbs = ref_bytes / (ms / 1000.0) / 1e9
        return {"bound": "l2", "kernel": kernel, "achieved": l2["gbs"] if l2 else None, "peak": l2_peak,
                "unit": "GB/s", "frac": l2["frac"] if l2 else None, "traffic": traffic, "avg_launch_ms": ms,
                "l2_request_bytes_per_launch": l2["bytes_per_launch"] if l2 else None,
                "valu_issue": valu_issue(kernel, ms, ceil) if ordered else None,
                "hbm": {"peak": HBM_PEAK_GBS, "traffic_gbs": traffic / (ms / 1000.0) / 1e9 if traffic else None,
                        "frac_traffic": traffic / (ms / 1000.0) / 1e9 / HBM_PEAK_GBS if traffic else None},
                # 32 B a BLAS node, 48 B a TLAS node, 64 B an instance entry, 36 B a triangle test:
                # the reference's layout; this walk reads the TLAS and instances from LDS and
                # re-reads nodes from L1, so these bytes are a work count, not traffic
                "reference_layout_bytes": {"per_launch": ref_bytes, "gbs": ref_gbs},
                "sources": ["profiles/roofline_ceilings.md", "profiles/walk_instanced_pmc.md"]}

    blend_ms = float(np.mean(stage["blend"]))
    T_irr, T_d = 36, w.depth_oct * w.depth_oct
    blend_flops = 2.0 * w.probes * w.rays * (5 * T_irr + 2 * T_d)
    # the FP32 ray records k_resolve writes (16 B a ray), and the previous and new interior texels
    blend_bytes = w.probes * w.rays * 16 + w.probes * (T_irr * 32 + T_d * 16)
    blend_kernel = {0: "k_blend_fp32", 1: "k_blend_mfma<1u,13,false>", 2: "k_blend_mfma<2u,13,false>",
                    3: "k_blend_mfma<3u,13,false>"}[prec]
    blend_traffic = measured_traffic(blend_kernel)

    fp64 = None if args.no_check else fp64_check(w, states, w.depth_oct, sample_ranges(w.probes))
    blend_mfma = mfma_util(blend_kernel) if prec else None
    sampling = sample_leg(f, w)
    f.close()
    adj = None if args.no_adjoint else adjoint_leg()
    dyn = None if args.no_dynamic else dynamic_leg(frames=args.dynamic_frames)
    bat = None if args.no_batch else batch_leg(frames=args.batch_frames)
    lights = None if args.no_lights else lights_leg()
    room = None if args.no_room else room_leg()
    reflect = None if args.no_reflect else reflect_leg()
    cpu = None if args.no_cpu else cpu_baseline(w)

    out = {
        "metric": METRIC, "value": value, "unit": "rays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": f"f32 traversal, {args.precision} blend", "data": "synthetic",
        "config": {"workload": "instanced: 8 meshes in 256 instances, 261514 triangles, 16 point lights, "
                               "32x16x24 probes x 256 rays, irr 8x8, depth 16x16, 2 bounces",
                   "probes": w.probes, "rays_per_probe": w.rays, "triangles": w.scene.meta.get("instanced_triangles"),
                   "unique_triangles": w.scene.meta.get("unique_triangles"), "lights": 16,
                   "blend_precision": args.precision, "traversal": args.traversal},
        "ms_per_update": ms_step,
        "stage_ms": {k: float(np.mean(v)) for k, v in stage.items()},
        "shadow_rays_per_update": int(cs.shadow_rays),
        "total_rays_per_s": (rays + cs.shadow_rays) / (ms_step / 1000.0),
        "irradiance_rel_err_vs_fp64": fp64,
        # roofline of record: the shadow walk's L2 read bytes against the measured L2 read
        # ceiling (its nodes come from L2 and L1; HBM carries ~1 % of its reads); VALU issue
        # and HBM beside it
        "roofline": dict(walk_roofline(shadow_kernel, sw_bytes, sw_ms),
                         # top-level steps: TLAS nodes, or candidate-list entries with the lists
                         top_level="candidate lists" if lists_on else "tlas",
                         counts={"blas_nodes": int(cs.shadow_blas_nodes), "tlas_nodes": int(cs.shadow_tlas_nodes),
                                 "tri_tests": int(cs.shadow_tri_tests), "inst_visits": int(cs.shadow_inst_visits),
                                 "rays": int(cs.shadow_rays)}),
        "primary_walk_roofline": walk_roofline("k_walk<0,false,true,PrimarySrc,false,256>", pw_bytes, pw_ms),
        "trace": {"algorithmic_bytes": bytes_alg, "ms": trace_ms, "gbs": bytes_alg / (trace_ms / 1000.0) / 1e9,
                  "counts": {"blas_nodes": int(cs.blas_nodes), "tlas_nodes": int(cs.tlas_nodes),
                             "tri_tests": int(cs.tri_tests), "inst_visits": int(cs.inst_visits),
                             "primary_hits": hits},
                  "shadow_walk_phases": shadow_phases(cs)},
        "blend": {"kernel": blend_kernel, "ms": blend_ms, "tflops": blend_flops / (blend_ms / 1000.0) / 1e12,
                  "gbs": blend_bytes / (blend_ms / 1000.0) / 1e9, "flops": blend_flops, "bytes": blend_bytes,
                  "hbm_traffic": blend_traffic,
                  "hbm_traffic_frac": blend_traffic / (blend_ms / 1000.0) / 1e9 / HBM_PEAK_GBS if blend_traffic else None,
                  "mfma_busy": blend_mfma},
        "room": room,
        "fwd_plus_adjoint_ms": adj["fwd_plus_adjoint_ms"] if adj else None,
        "adjoint": adj,
        "dynamic": dyn,
        "batch": bat,
        "lights": lights,
        "reflections": reflect,
        "receiver_sampling": sampling,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
