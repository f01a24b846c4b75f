"""Host-side probe field over the C ABI (the Python mirror of createProbeField's
solve path, js/gi_probes.js:945-4847, for tests and benchmarks).

    f = ProbeField(res=(16,16,16), grid_min=..., grid_size=..., rays=64)
    f.set_scene(scene)
    f.tick(frame_jitter=j, hysteresis=0.97)   # one full-field update
    irr = f.read_irr()

Everything runs through libsbgi.so; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


def clamp_rays(n) -> int:
    """setRays: round to a multiple of 16 within [32, 256] (js/gi_probes.js:4589-4594)."""
    try:
        v = int(round(float(n) / 16.0)) * 16
    except (TypeError, ValueError):
        v = 0
    if v == 0:
        v = 64
    return max(32, min(256, v))


class ProbeField:
    def __init__(self, res, grid_min, grid_size, rays: int = 64, depth_oct: int = 6,
                 precision: int = N.BLEND_FP32, keep_hit_ids: bool = False, device: int = 0, cascades=None,
                 traversal: int = N.TRAV_MT_COMPAT, clustered: bool = False, reflections: int = N.REFLECT_OFF,
                 graph: bool = True, shadow_lists: bool = True):
        self.L = N.lib()
        d = N.FieldDesc()
        d.device = device
        d.rays_per_probe = rays
        casc = cascades or [(grid_min, grid_size, res)]
        d.cascade_count = len(casc)
        d.depth_oct_res = depth_oct
        d.blend_precision = precision
        d.traversal = traversal
        d.flags = ((N.FIELD_KEEP_HIT_IDS if keep_hit_ids else 0) | (N.FIELD_CLUSTERED if clustered else 0) |
                   (0 if graph else N.FIELD_NO_GRAPH) | (0 if shadow_lists else N.FIELD_NO_SHADOW_LISTS))
        d.reflections = reflections
        for i, (gm, gs, rs) in enumerate(casc):
            for a in range(3):
                d.cascades[i].grid_min[a] = gm[a]
                d.cascades[i].grid_size[a] = gs[a]
                d.cascades[i].res[a] = rs[a]
        h = C.c_void_p()
        N.check(self.L.sbgi_create(C.byref(d), C.byref(h)), None)
        self.h = h
        self.desc = d
        self.rays = rays
        self.depth_oct = depth_oct
        self.cascades = casc
        self.needs_clear = [True] * len(casc)
        self.last_ticket = 0

    def close(self):
        if getattr(self, "h", None):
            self.L.sbgi_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, code):
        return N.check(code, self.h)

    def probes(self, cascade: int = 0) -> int:
        r = self.cascades[cascade][2]
        return r[0] * r[1] * r[2]

    def set_scene(self, scene):
        self._scene = scene  # keep the arrays alive for the call
        self._chk(self.L.sbgi_set_scene(self.h, C.byref(scene.view())))

    def tick(self, cascade: int = 0, probe_offset: int = 0, updated: int = 0, wait: bool = True, **params) -> int:
        p = N.default_tick_params(**params)
        p.cascade = cascade
        p.probe_offset = probe_offset
        p.updated_count = updated
        p.needs_clear = 1 if self.needs_clear[cascade] else 0
        self.needs_clear[cascade] = False
        t = C.c_uint64()
        self._chk(self.L.sbgi_tick(self.h, C.byref(p), C.byref(t)))
        self.last_ticket = t.value
        if wait:
            self.wait(t.value)
        return t.value

    def wait(self, ticket=None, timeout_ms: int = -1):
        self._chk(self.L.sbgi_wait(self.h, self.last_ticket if ticket is None else ticket, timeout_ms))

    def poll(self, ticket) -> bool:
        return self._chk(self.L.sbgi_poll(self.h, ticket)) == 1

    def stats(self) -> N.Stats:
        s = N.Stats()
        self._chk(self.L.sbgi_get_stats(self.h, C.byref(s)))
        return s

    def set_counting(self, on: bool):
        self._chk(self.L.sbgi_set_counting(self.h, int(on)))

    def set_stage_timing(self, on: bool):
        """Record per-stage HIP events in graph-replayed ticks (stats().ms_*); off by default."""
        self._chk(self.L.sbgi_set_stage_timing(self.h, int(on)))

    def nbytes(self, buf: int, cascade: int = 0) -> int:
        return int(self.L.sbgi_buffer_bytes(self.h, buf, cascade))

    def read(self, buf: int, dtype, cascade: int = 0) -> np.ndarray:
        n = self.nbytes(buf, cascade)
        out = np.empty(n // np.dtype(dtype).itemsize, dtype)
        self._chk(self.L.sbgi_read(self.h, buf, cascade, N.ptr(out), n))
        return out

    def write(self, buf: int, data: np.ndarray, cascade: int = 0):
        data = np.ascontiguousarray(data)
        self._chk(self.L.sbgi_write(self.h, buf, cascade, N.ptr(data), data.nbytes))

    def read_irr(self, cascade=0):
        return self.read(N.BUF_IRR_STATE, np.float32, cascade).reshape(-1, 64, 4)

    def read_depth(self, cascade=0):
        t = 64 if self.depth_oct == 6 else (self.depth_oct + 2) ** 2
        return self.read(N.BUF_DEPTH_STATE, np.float32, cascade).reshape(-1, t, 2)

    def read_atlas(self, depth=False, cascade=0):
        raw = self.read(N.BUF_DEPTH_ATLAS if depth else N.BUF_IRR_ATLAS, np.uint16, cascade)
        res = self.cascades[cascade][2]
        tile = 8 if (not depth or self.depth_oct == 6) else self.depth_oct + 2
        return raw.view(np.float16).astype(np.float32).reshape(res[1] * res[2] * tile, res[0] * tile, 4)

    def read_emitter_vis(self, cascade=0):
        """Retained emitter visibility [probe, 16]."""
        return self.read(N.BUF_EMITTER_VIS, np.float32, cascade).reshape(-1, N.MAX_EMITTERS)

    def read_light_grid(self):
        """Clustered light lists [cell, 17] of cascade 0."""
        return self.read(N.BUF_LIGHT_GRID, np.uint32).reshape(-1, N.LIGHTS_PER_CELL + 1)

    def read_rays(self):
        return self.read(N.BUF_RAY_DATA, np.float32).reshape(-1, self.rays, 4)

    def read_hit_ids(self):
        return self.read(N.BUF_HIT_IDS, np.int32).reshape(-1, self.rays, 2)

    # ---- device skinning (sbgi_set_skins / sbgi_skin) ----
    def set_skins(self, sets, n_bones: int):
        """sets: iterable of (vertex_start, vertex_count, blas_root, blas_end, joints u16[V,4],
        weights f32[V,4]); the rest pose is the vertexData last provided by the host."""
        keep, arr = [], (N.SkinSet * max(1, len(sets)))()
        for i, (vs, vc, br, be, j, w) in enumerate(sets):
            j = np.ascontiguousarray(j, np.uint16).reshape(-1, 4)
            w = np.ascontiguousarray(w, np.float32).reshape(-1, 4)
            keep += [j, w]
            arr[i] = N.SkinSet(vs, vc, br, be, j.ctypes.data, w.ctypes.data)
        self._chk(self.L.sbgi_set_skins(self.h, arr, len(sets), n_bones))

    def skin(self, bones, update_tlas: bool = True):
        """One skinning frame; bones: [n_bones, 3, 4] row-major world-from-rest matrices."""
        b = np.ascontiguousarray(bones, np.float32).reshape(-1, 12)
        self._chk(self.L.sbgi_skin(self.h, N.ptr(b), b.shape[0], int(update_tlas)))

    def edit_timing(self):
        """(skin ms, BLAS refit ms, instance bounds + TLAS ms) of the last scene edits."""
        a, b, c = C.c_float(), C.c_float(), C.c_float()
        self._chk(self.L.sbgi_edit_timing(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    # ---- walk schedule ----
    def walk_tuning(self) -> dict:
        t = N.WalkTuning()
        self._chk(self.L.sbgi_get_walk_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in N.WalkTuning._fields_}

    def set_walk_tuning(self, **kw):
        t = N.WalkTuning()
        self._chk(self.L.sbgi_get_walk_tuning(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        self._chk(self.L.sbgi_set_walk_tuning(self.h, C.byref(t)))

    def autotune(self, cascade: int = 0, **params):
        """Coordinate-descent search of the walk schedule on this scene; keeps the best.
        Returns (tuning dict, trace ms)."""
        p = N.default_tick_params(**params)
        p.cascade = cascade
        t, ms = N.WalkTuning(), C.c_float()
        self._chk(self.L.sbgi_autotune(self.h, C.byref(p), C.byref(t), C.byref(ms)))
        return {k: getattr(t, k) for k, _ in N.WalkTuning._fields_}, ms.value

    # ---- differentiable mode (sbgi_diff_*) ----
    def diff_param_count(self):
        """(albedo entries, lights) of the differentiable mode."""
        n, L = C.c_uint32(), C.c_uint32()
        self._chk(self.L.sbgi_diff_param_count(self.h, C.byref(n), C.byref(L)))
        return n.value, L.value

    def diff_albedo_map(self):
        """Per TLAS slot: (offset into the albedo array, first pool triangle)."""
        m = np.zeros((self._scene.inst_count, 2), np.int32)
        self._chk(self.L.sbgi_diff_albedo_map(self.h, N.ptr(m), self._scene.inst_count))
        return m

    def diff_forward(self, albedo, light_scale=None, **params):
        """Second-update interior irradiance (P, 36, 3) of the two-update differentiable solve."""
        p = N.default_tick_params(**params)
        alb = np.ascontiguousarray(albedo, np.float32)
        ls = None if light_scale is None else np.ascontiguousarray(light_scale, np.float32)
        out = np.zeros((self.probes(0), 36, 3), np.float32)
        self._keep = (alb, ls)
        self._chk(self.L.sbgi_diff_forward(self.h, C.byref(p), N.ptr(alb), None if ls is None else N.ptr(ls),
                                           N.ptr(out)))
        return out

    def diff_adjoint(self, grad_irr):
        """(d/d albedo (n, 3), d/d light scale (L,)) of sum(grad_irr * irr) at the last forward."""
        n, L = self.diff_param_count()
        g = np.ascontiguousarray(grad_irr, np.float32)
        ga = np.zeros((n, 3), np.float32)
        gl = np.zeros(max(L, 1), np.float32)
        self._chk(self.L.sbgi_diff_adjoint(self.h, N.ptr(g), N.ptr(ga), N.ptr(gl)))
        return ga, gl[:L]

    def diff_timing(self):
        """(forward ms, adjoint ms) of the last differentiable forward / adjoint (HIP events)."""
        a, b = C.c_float(), C.c_float()
        self._chk(self.L.sbgi_diff_timing(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def query(self, ro, rd, tmax=None, mode=0, noalloc=False):
        """Closest-hit (mode 0) or any-hit (mode 1) queries of the resident scene, through
        sbgi_query_rays or (noalloc) sbgi_query_rays_noalloc. Returns (t, tri, inst, blocked)
        arrays."""
        ro = np.asarray(ro, np.float32).reshape(-1, 3)
        rd = np.asarray(rd, np.float32).reshape(-1, 3)
        n = ro.shape[0]
        rec = np.zeros((n, 8), np.float32)
        rec[:, 0:3] = ro
        rec[:, 3] = 1e30 if tmax is None else tmax
        rec[:, 4:7] = rd
        rec[:, 7] = mode
        out = np.zeros((n, 4), np.float32)
        fn = self.L.sbgi_query_rays_noalloc if noalloc else self.L.sbgi_query_rays
        self._chk(fn(self.h, N.ptr(rec), n, N.ptr(out)))
        return out[:, 0], out[:, 1].view(np.int32), out[:, 2].view(np.int32), out[:, 3] > 0.5

    def sample(self, pos, nvis, ndir=None, **params):
        """Irradiance of the field at n points (sbgi_sample_irradiance): pos, nvis (and
        ndir) are (n, 3); returns ((n, 3) float32, device ms)."""
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        nvis = np.ascontiguousarray(nvis, np.float32).reshape(-1, 3)
        nd = None if ndir is None else np.ascontiguousarray(ndir, np.float32).reshape(-1, 3)
        out = np.empty_like(pos)
        ms = C.c_float()
        p = N.default_sample_params(**params)
        self._chk(self.L.sbgi_sample_irradiance(self.h, C.byref(p), N.ptr(pos), N.ptr(nvis), N.ptr(nd), pos.shape[0],
                                                N.ptr(out), C.byref(ms)))
        return out, ms.value

    def update_ranges(self, buf: int, ranges, src: np.ndarray):
        r = np.ascontiguousarray(np.asarray(ranges, np.uint32).reshape(-1))
        src = np.ascontiguousarray(src)
        self._chk(self.L.sbgi_update_ranges(self.h, buf, N.ptr(r), r.size // 2, N.ptr(src)))

    def refit(self, blas_ranges, refit_tlas: bool = True):
        r = np.ascontiguousarray(np.asarray(blas_ranges, np.uint32).reshape(-1))
        self._chk(self.L.sbgi_refit(self.h, N.ptr(r) if r.size else None, r.size // 2, int(refit_tlas)))

    def classify(self, cascade: int = 0):
        """Probe classification of one cascade (sbgi_classify): writes every probe's meta
        {state, relocation offset}, read back with read(BUF_META)."""
        self._chk(self.L.sbgi_classify(self.h, cascade))

    def rebuild_blas(self, blas_ranges):
        """Rebuild the BLASes of (root, end) node ranges on the device (sbgi_rebuild_blas)."""
        r = np.ascontiguousarray(np.asarray(blas_ranges, np.uint32).reshape(-1))
        self._chk(self.L.sbgi_rebuild_blas(self.h, N.ptr(r), r.size // 2))

    def set_instance_transforms(self, slots, worlds, active=None):
        s = np.ascontiguousarray(np.asarray(slots, np.uint32))
        w = np.ascontiguousarray(np.stack([np.asarray(m, np.float64).T.reshape(-1) for m in worlds]))
        a = None if active is None else np.ascontiguousarray(np.asarray(active, np.uint8))
        self._chk(self.L.sbgi_set_instance_transforms(self.h, N.ptr(s), s.size, N.ptr(w), N.ptr(a)))


class Batch:
    """Batched solves of several ProbeFields (sbgi_batch_*): one frame updates every cascade
    of every field with one persistent closest-hit walk and one persistent shadow walk over
    all (field, cascade) segments, replayed from a hipGraph while the arguments repeat."""

    def __init__(self, fields):
        self.L = N.lib()
        self.fields = list(fields)
        arr = (C.c_void_p * len(self.fields))(*[f.h.value for f in self.fields])
        h = C.c_void_p()
        code = self.L.sbgi_batch_create(arr, len(self.fields), C.byref(h))
        if code < 0:
            raise N.SbgiError(code, "sbgi_batch_create failed")
        self.h = h

    def _chk(self, code):
        if code < 0:
            raise N.SbgiError(code, (self.L.sbgi_batch_last_error(self.h) or b"").decode(errors="replace"))
        return code

    def _params(self, params):
        arr = (N.TickParams * len(self.fields))()
        for i, (f, kw) in enumerate(zip(self.fields, params)):
            arr[i] = N.default_tick_params(**kw)
            arr[i].needs_clear = 1 if f.needs_clear[0] else 0
        return arr

    def tick(self, params, wait: bool = True) -> int:
        """params: one keyword dict of tick parameters per field."""
        arr = self._params(params)
        t = C.c_uint64()
        self._chk(self.L.sbgi_batch_tick(self.h, arr, C.byref(t)))
        for f in self.fields:
            f.needs_clear = [False] * len(f.needs_clear)
        if wait:
            self.wait()
        return t.value

    def wait(self):
        self._chk(self.L.sbgi_batch_wait(self.h))

    def timing(self):
        """(device ms of the last frame, whether it replayed the captured graph)."""
        ms, rp = C.c_float(), C.c_int32()
        self._chk(self.L.sbgi_batch_timing(self.h, C.byref(ms), C.byref(rp)))
        return ms.value, bool(rp.value)

    def set_walk_block(self, block: int):
        self._chk(self.L.sbgi_batch_set_walk_block(self.h, block))

    @property
    def walk_waves(self) -> int:
        """Resident waves per CU of the batched walks (0: as many as fit)."""
        w = C.c_uint32()
        self._chk(self.L.sbgi_batch_get_walk_waves(self.h, C.byref(w)))
        return w.value

    @walk_waves.setter
    def walk_waves(self, waves: int):
        self._chk(self.L.sbgi_batch_set_walk_waves(self.h, waves))

    def autotune(self, params):
        arr = self._params(params)
        for i in range(len(self.fields)):
            arr[i].needs_clear = 0
        t, blk, ms = N.WalkTuning(), C.c_uint32(), C.c_float()
        self._chk(self.L.sbgi_batch_autotune(self.h, arr, C.byref(t), C.byref(blk), C.byref(ms)))
        return {k: getattr(t, k) for k, _ in N.WalkTuning._fields_}, blk.value, ms.value

    def close(self):
        if getattr(self, "h", None):
            self.L.sbgi_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
