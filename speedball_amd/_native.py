"""ctypes binding of libsbgi.so (include/sbgi.h).

The library is loaded from this package directory. There is no fallback: when the
library is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsbgi.so")

SBGI_OK = 0
SBGI_E_INVALID = -1
SBGI_E_LAYOUT = -2
SBGI_E_CAPACITY = -3
SBGI_E_DEVICE = -4
SBGI_E_STATE = -5
SBGI_E_TIMEOUT = -6
SBGI_E_NOMEM = -7

BLEND_FP32, BLEND_BF16, BLEND_BF16X3, BLEND_FP16 = 0, 1, 2, 3
TRAV_MT_COMPAT, TRAV_MT_ORDERED, TRAV_WATERTIGHT = 0, 1, 2
FIELD_KEEP_HIT_IDS = 1
FIELD_CLUSTERED = 2
FIELD_NO_GRAPH = 4
FIELD_NO_SHADOW_LISTS = 8
MAX_EMITTERS = 16
LIGHTS_PER_CELL = 16
EMITTER_RETENTION = 0.8

BUF_BVH_NODES, BUF_TRI_INDEX, BUF_VERTEX_DATA, BUF_TRI_MATERIAL, BUF_MATERIALS, BUF_LIGHTS = 0, 1, 2, 3, 4, 5
BUF_IRR_STATE, BUF_DEPTH_STATE, BUF_META, BUF_IRR_ATLAS, BUF_DEPTH_ATLAS = 10, 11, 12, 13, 14
BUF_RAY_DATA, BUF_HIT_IDS, BUF_NODES_REFIT, BUF_TRI_PACKED, BUF_INST_BOUNDS, BUF_TLAS_SAH = 15, 16, 17, 18, 19, 20
BUF_EMITTER_VIS, BUF_LIGHT_GRID = 21, 22
BUF_ROUGH_STATE, BUF_ROUGH_ATLAS, BUF_GLOSSY_STATE, BUF_GLOSSY_WEIGHT, BUF_GLOSSY_ATLAS = 23, 24, 25, 26, 27
BUF_TLAS_OCT = 28
BUF_SHADOW_VIS = 29
BUF_SHADOW_PROJ, BUF_SHADOW_CELLS, BUF_SHADOW_LIST = 30, 31, 32
REFLECT_OFF, REFLECT_ROUGH, REFLECT_HIGH, REFLECT_ULTRA = 0, 1, 2, 3

# every symbol include/sbgi.h declares
EXPORTS = [
    "sbgi_tick_params_default", "sbgi_version", "sbgi_create", "sbgi_set_scene", "sbgi_update_ranges",
    "sbgi_refit", "sbgi_set_instance_transforms", "sbgi_tick", "sbgi_wait", "sbgi_poll", "sbgi_read",
    "sbgi_write", "sbgi_buffer_bytes", "sbgi_query_rays", "sbgi_query_rays_noalloc", "sbgi_get_stats", "sbgi_set_counting", "sbgi_set_stage_timing", "sbgi_last_error",
    "sbgi_destroy", "sbgi_sb_create", "sbgi_sb_add_material", "sbgi_sb_add_mesh", "sbgi_sb_add_instance",
    "sbgi_sb_add_light", "sbgi_sb_finalize", "sbgi_sb_tlas_order", "sbgi_sb_set_transform",
    "sbgi_sb_last_error", "sbgi_sb_destroy", "sbgi_diff_param_count", "sbgi_diff_albedo_map", "sbgi_diff_forward",
    "sbgi_diff_adjoint", "sbgi_diff_timing", "sbgi_get_walk_tuning", "sbgi_set_walk_tuning", "sbgi_autotune",
    "sbgi_set_skins", "sbgi_skin", "sbgi_edit_timing", "sbgi_sb_mesh_ranges",
    "sbgi_batch_create", "sbgi_batch_tick", "sbgi_batch_wait", "sbgi_batch_timing", "sbgi_batch_set_walk_block",
    "sbgi_batch_set_walk_waves", "sbgi_batch_get_walk_waves", "sbgi_batch_autotune", "sbgi_batch_last_error", "sbgi_batch_destroy", "sbgi_sb_add_materials",
    "sbgi_rebuild_blas", "sbgi_sample_params_default", "sbgi_sample_irradiance", "sbgi_classify",
]


class CascadeDesc(C.Structure):
    _fields_ = [("grid_min", C.c_float * 3), ("grid_size", C.c_float * 3), ("res", C.c_uint32 * 3),
                ("pad", C.c_uint32)]


class FieldDesc(C.Structure):
    _fields_ = [("device", C.c_int32), ("rays_per_probe", C.c_uint32), ("cascade_count", C.c_uint32),
                ("depth_oct_res", C.c_uint32), ("blend_precision", C.c_uint32), ("traversal", C.c_uint32),
                ("flags", C.c_uint32), ("reflections", C.c_uint32), ("cascades", CascadeDesc * 4)]


class SceneView(C.Structure):
    _fields_ = [("bvh_nodes", C.c_void_p), ("node_count", C.c_uint32), ("tri_count", C.c_uint32),
                ("tri_index", C.c_void_p), ("vertex_data", C.c_void_p), ("vertex_count", C.c_uint32),
                ("materials_len", C.c_uint32), ("tri_material", C.c_void_p), ("materials", C.c_void_p),
                ("inst_base", C.c_uint32), ("inst_count", C.c_uint32), ("tlas_base", C.c_uint32),
                ("tlas_node_count", C.c_uint32), ("lights", C.c_void_p), ("light_count", C.c_uint32),
                ("emitter_base", C.c_uint32), ("emitter_count", C.c_uint32), ("pad", C.c_uint32)]


TICK_FLOATS = [
    "frame_jitter", "hysteresis", "hysteresis_exponent", "depth_sharpness", "radiance_clamp",
    "classify_strength", "filter_strength", "filter_smooth", "temp_change_sigma1", "temp_change_h_drop",
    "temp_clamp_sigma", "nir_gate", "nir_gain", "sky_intensity",
]
TICK_DEBUG = [
    "dbg_trace_bias", "dbg_ray_eps", "dbg_direct", "dbg_emissive", "dbg_albedo", "dbg_bounce",
    "dbg_cosine_power", "dbg_noise_h_boost", "dbg_change_sigma0", "dbg_min_change_h", "dbg_var_eps",
    "dbg_var_rel", "dbg_depth_history", "dbg_filter_k", "dbg_filter_rel", "dbg_filter_eps",
]


class TickParams(C.Structure):
    _fields_ = ([("cascade", C.c_uint32), ("probe_offset", C.c_uint32), ("updated_count", C.c_uint32),
                 ("needs_clear", C.c_uint32)]
                + [(n, C.c_float) for n in TICK_FLOATS]
                + [("sky_sh", C.c_float * 27)]
                + [(n, C.c_float) for n in TICK_DEBUG]
                + [("emitter_vis_seed", C.c_float), ("sky_configured", C.c_float),
                   ("reflection_sky_fallback", C.c_float), ("glossy_phase", C.c_uint32), ("pad", C.c_uint32)])


class Stats(C.Structure):
    _fields_ = [("ms_raydir", C.c_float), ("ms_trace", C.c_float), ("ms_blend", C.c_float),
                ("ms_upload", C.c_float), ("ms_total", C.c_float), ("rays", C.c_uint32),
                ("updated", C.c_uint32), ("shadow_rays", C.c_uint64), ("blas_nodes", C.c_uint64),
                ("tlas_nodes", C.c_uint64), ("tri_tests", C.c_uint64), ("inst_visits", C.c_uint64),
                ("ms_walk_primary", C.c_float), ("ms_emit", C.c_float), ("ms_walk_shadow", C.c_float),
                ("ms_resolve", C.c_float), ("shadow_blas_nodes", C.c_uint64), ("shadow_tlas_nodes", C.c_uint64),
                ("shadow_tri_tests", C.c_uint64), ("shadow_inst_visits", C.c_uint64),
                ("walk_waves", C.c_uint64 * 4), ("shadow_waves", C.c_uint64 * 4), ("ms_emitter_vis", C.c_float),
                ("tick_replayed", C.c_uint32), ("stage_timed", C.c_uint32)]


class SkinSet(C.Structure):
    _fields_ = [("vertex_start", C.c_uint32), ("vertex_count", C.c_uint32), ("blas_root", C.c_uint32),
                ("blas_end", C.c_uint32), ("joints", C.c_void_p), ("weights", C.c_void_p)]


class WalkTuning(C.Structure):
    _fields_ = [("tri_ratio", C.c_uint32), ("inst_ratio", C.c_uint32), ("refill_min", C.c_uint32),
                ("fetch_chunk", C.c_uint32), ("shadow_lists", C.c_uint32), ("shadow_copies", C.c_uint32)]


class SampleParams(C.Structure):
    _fields_ = [("normal_bias", C.c_float), ("cheby_strength", C.c_float), ("classify_strength", C.c_float),
                ("border_band", C.c_float), ("cheby_bias", C.c_float * 4), ("cascades", C.c_uint32),
                ("pad", C.c_uint32)]


def default_sample_params(**kw) -> SampleParams:
    p = SampleParams()
    lib().sbgi_sample_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "cheby_bias":
            for i, x in enumerate(v):
                p.cheby_bias[i] = float(x)
        else:
            setattr(p, k, v)
    return p


class SbgiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sbgi error {code}: {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    """Load libsbgi.so; raises when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `python -m speedball_amd.build` "
                           "(the HIP extension is required; there is no fallback path)")
    # PyTorch ships its own libamdhip64.so whose SONAME (libamdhip64.so.7) equals the
    # system runtime's. Loading torch first makes its copy THE process's HIP runtime, which
    # libsbgi.so then binds to by SONAME; loading libsbgi.so first would put two HIP
    # runtimes in one process and torch would see no device.
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    L = C.CDLL(LIB_PATH)
    vp, u32, i32, f32p = C.c_void_p, C.c_uint32, C.c_int32, C.c_void_p
    sig = {
        "sbgi_tick_params_default": (None, [C.POINTER(TickParams)]),
        "sbgi_version": (C.c_char_p, []),
        "sbgi_create": (i32, [C.POINTER(FieldDesc), C.POINTER(vp)]),
        "sbgi_set_scene": (i32, [vp, C.POINTER(SceneView)]),
        "sbgi_update_ranges": (i32, [vp, i32, vp, u32, vp]),
        "sbgi_refit": (i32, [vp, vp, u32, i32]),
        "sbgi_set_instance_transforms": (i32, [vp, vp, u32, vp, vp]),
        "sbgi_tick": (i32, [vp, C.POINTER(TickParams), C.POINTER(C.c_uint64)]),
        "sbgi_wait": (i32, [vp, C.c_uint64, i32]),
        "sbgi_poll": (i32, [vp, C.c_uint64]),
        "sbgi_read": (i32, [vp, i32, u32, vp, C.c_size_t]),
        "sbgi_write": (i32, [vp, i32, u32, vp, C.c_size_t]),
        "sbgi_buffer_bytes": (C.c_size_t, [vp, i32, u32]),
        "sbgi_get_stats": (i32, [vp, C.POINTER(Stats)]),
        "sbgi_set_counting": (i32, [vp, i32]),
        "sbgi_set_stage_timing": (i32, [vp, i32]),
        "sbgi_query_rays": (i32, [vp, vp, u32, vp]),
        "sbgi_query_rays_noalloc": (i32, [vp, vp, u32, vp]),
        "sbgi_sample_params_default": (None, [C.POINTER(SampleParams)]),
        "sbgi_sample_irradiance": (i32, [vp, C.POINTER(SampleParams), vp, vp, vp, u32, vp, vp]),
        "sbgi_diff_param_count": (i32, [vp, C.POINTER(u32), C.POINTER(u32)]),
        "sbgi_diff_albedo_map": (i32, [vp, vp, u32]),
        "sbgi_diff_forward": (i32, [vp, C.POINTER(TickParams), vp, vp, vp]),
        "sbgi_diff_adjoint": (i32, [vp, vp, vp, vp]),
        "sbgi_diff_timing": (i32, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "sbgi_get_walk_tuning": (i32, [vp, C.POINTER(WalkTuning)]),
        "sbgi_set_walk_tuning": (i32, [vp, C.POINTER(WalkTuning)]),
        "sbgi_autotune": (i32, [vp, C.POINTER(TickParams), C.POINTER(WalkTuning), C.POINTER(C.c_float)]),
        "sbgi_set_skins": (i32, [vp, C.POINTER(SkinSet), u32, u32]),
        "sbgi_rebuild_blas": (i32, [vp, vp, u32]),
        "sbgi_classify": (i32, [vp, u32]),
        "sbgi_skin": (i32, [vp, vp, u32, i32]),
        "sbgi_edit_timing": (i32, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "sbgi_batch_create": (i32, [C.POINTER(vp), u32, C.POINTER(vp)]),
        "sbgi_batch_tick": (i32, [vp, C.POINTER(TickParams), C.POINTER(C.c_uint64)]),
        "sbgi_batch_wait": (i32, [vp]),
        "sbgi_batch_timing": (i32, [vp, C.POINTER(C.c_float), C.POINTER(i32)]),
        "sbgi_batch_set_walk_block": (i32, [vp, u32]),
        "sbgi_batch_set_walk_waves": (i32, [vp, u32]),
        "sbgi_batch_get_walk_waves": (i32, [vp, C.POINTER(u32)]),
        "sbgi_batch_autotune": (i32, [vp, C.POINTER(TickParams), C.POINTER(WalkTuning), C.POINTER(u32),
                                      C.POINTER(C.c_float)]),
        "sbgi_batch_last_error": (C.c_char_p, [vp]),
        "sbgi_batch_destroy": (i32, [vp]),
        "sbgi_last_error": (C.c_char_p, [vp]),
        "sbgi_destroy": (i32, [vp]),
        "sbgi_sb_create": (i32, [C.POINTER(vp)]),
        "sbgi_sb_add_material": (i32, [vp, f32p, C.POINTER(u32)]),
        "sbgi_sb_add_materials": (i32, [vp, vp, u32, C.POINTER(u32)]),
        "sbgi_sb_add_mesh": (i32, [vp, f32p, u32, f32p, f32p, vp, u32, vp, C.POINTER(u32)]),
        "sbgi_sb_add_instance": (i32, [vp, u32, vp, i32, C.POINTER(u32)]),
        "sbgi_sb_add_light": (i32, [vp, f32p]),
        "sbgi_sb_finalize": (i32, [vp, C.POINTER(SceneView)]),
        "sbgi_sb_tlas_order": (i32, [vp, vp, u32]),
        "sbgi_sb_set_transform": (i32, [vp, u32, vp, C.POINTER(u32)]),
        "sbgi_sb_mesh_ranges": (i32, [vp, u32, vp]),
        "sbgi_sb_last_error": (C.c_char_p, [vp]),
        "sbgi_sb_destroy": (i32, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(code: int, handle=None, builder: bool = False) -> int:
    if code < 0:
        L = lib()
        msg = (L.sbgi_sb_last_error(handle) if builder else L.sbgi_last_error(handle)) or b""
        raise SbgiError(code, msg.decode(errors="replace"))
    return code


def default_tick_params(**kw) -> TickParams:
    p = TickParams()
    lib().sbgi_tick_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "sky_sh":
            for i, x in enumerate(v):
                p.sky_sh[i] = float(x)
        else:
            setattr(p, k, v)
    return p


def ptr(a) -> C.c_void_p:
    """Data pointer of a C-contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return C.c_void_p(a.ctypes.data)
