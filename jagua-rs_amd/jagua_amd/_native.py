"""ctypes declarations of include/jagua_amd.h (libjagua_amd.so) and of the synthetic-workload
generator (libjagua_synth.so).  Loading fails loudly when the native library is missing:
there is no Python fallback for any query."""
import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(os.path.dirname(_PKG), "lib")

c_f32p = C.POINTER(C.c_float)
c_u8p = C.POINTER(C.c_uint8)
c_u32p = C.POINTER(C.c_uint32)
c_u64p = C.POINTER(C.c_uint64)
c_vp = C.c_void_p


class SurrogateConfig(C.Structure):
    _fields_ = [("n_pole_limits_n", C.c_uint32 * 3), ("n_pole_limits_cov", C.c_float * 3),
                ("n_ff_poles", C.c_uint32), ("n_ff_piers", C.c_uint32)]


class CDEConfigC(C.Structure):
    _fields_ = [("quadtree_depth", C.c_uint32), ("cd_threshold", C.c_uint32),
                ("item_surrogate_config", SurrogateConfig)]


class HazardEntityC(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("_pad", C.c_uint32), ("a", C.c_uint64), ("b", C.c_uint64),
                ("dt", C.c_float * 3), ("_pad2", C.c_uint32)]


class HazKeyC(C.Structure):
    _fields_ = [("idx", C.c_uint32), ("version", C.c_uint32)]


class FilterC(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("n_keys", C.c_uint32), ("keys", C.POINTER(HazKeyC)),
                ("seed", C.c_uint32), ("threshold", C.c_uint32)]


class SPolygonInfoC(C.Structure):
    _fields_ = [("n_vertices", C.c_uint32), ("n_poles", C.c_uint32), ("n_piers", C.c_uint32),
                ("n_ff_poles", C.c_uint32), ("area", C.c_float), ("diameter", C.c_float),
                ("bbox", C.c_float * 4), ("poi", C.c_float * 3), ("convex_hull_area", C.c_float)]


class CDEStatsC(C.Structure):
    _fields_ = [("n_nodes", C.c_uint64), ("n_entries", C.c_uint64), ("n_edge_refs", C.c_uint64),
                ("n_hazard_vertices", C.c_uint64), ("n_slots", C.c_uint64), ("device_bytes", C.c_uint64),
                ("flatten_ms", C.c_double), ("upload_ms", C.c_double), ("patch_ms", C.c_double),
                ("patch_bytes", C.c_uint64), ("n_patches", C.c_uint64), ("n_full_uploads", C.c_uint64)]


class QuerySetC(C.Structure):
    _fields_ = [("cde", c_vp), ("item", c_vp), ("first", C.c_uint64), ("count", C.c_uint64),
                ("filter", FilterC)]


class PlanConfigC(C.Structure):
    _fields_ = [("chunk", C.c_uint32), ("workgroups", C.c_uint32), ("lanes_per_query", C.c_uint32),
                ("flags", C.c_uint32), ("stages", C.c_uint32)]


class SearchConfigC(C.Structure):
    _fields_ = [("samples_per_round", C.c_uint32), ("uniform_rounds", C.c_uint32), ("local_rounds", C.c_uint32),
                ("pad", C.c_uint32), ("seed", C.c_uint64), ("ls_sd_transl", C.c_float * 2), ("ls_sd_rot", C.c_float * 2)]


class SearchItemC(C.Structure):
    _fields_ = [("cde", c_vp), ("item", c_vp), ("rotations", c_f32p), ("n_rotations", C.c_uint32),
                ("continuous", C.c_uint32), ("filter", FilterC)]


class SearchResultC(C.Structure):
    _fields_ = [("found", C.c_uint32), ("rotation", C.c_float), ("tx", C.c_float), ("ty", C.c_float),
                ("cost", C.c_float), ("x_max", C.c_float), ("y_max", C.c_float), ("samples", C.c_uint32)]


# name -> (restype, argtypes)
_SIGS = {
    "jg_spolygon_new": (C.c_int32, [c_f32p, C.c_uint32, C.POINTER(c_vp)]),
    "jg_spolygon_generate_surrogate": (C.c_int32, [c_vp, C.POINTER(SurrogateConfig)]),
    "jg_spolygon_transform_clone": (C.c_int32, [c_vp, c_f32p, C.POINTER(c_vp)]),
    "jg_spolygon_new_batch": (C.c_int32, [c_f32p, C.POINTER(C.c_uint32), C.c_uint32, C.POINTER(SurrogateConfig),
                                          C.POINTER(c_vp), C.POINTER(C.c_uint32)]),
    "jg_spolygon_free": (None, [c_vp]),
    "jg_spolygon_get_info": (C.c_int32, [c_vp, C.POINTER(SPolygonInfoC)]),
    "jg_spolygon_get_vertices": (C.c_int32, [c_vp, c_f32p]),
    "jg_spolygon_get_poles": (C.c_int32, [c_vp, c_f32p]),
    "jg_spolygon_get_piers": (C.c_int32, [c_vp, c_f32p]),
    "jg_transforms_from_dt": (C.c_int32, [c_f32p, C.c_uint64, c_f32p]),
    "jg_dt_chain": (C.c_int32, [c_f32p, C.c_uint32, c_f32p, c_f32p]),
    "jg_spolygon_centroid": (C.c_int32, [c_vp, c_f32p]),
    "jg_cde_new": (C.c_int32, [c_f32p, C.POINTER(CDEConfigC), C.POINTER(HazardEntityC), C.POINTER(c_vp),
                               c_u8p, C.c_uint32, C.POINTER(c_vp)]),
    "jg_cde_free": (None, [c_vp]),
    "jg_cde_register_hazard": (C.c_int32, [c_vp, C.POINTER(HazardEntityC), c_vp, C.c_uint8, C.POINTER(HazKeyC)]),
    "jg_cde_deregister_hazard_by_entity": (C.c_int32, [c_vp, C.POINTER(HazardEntityC)]),
    "jg_cde_deregister_hazard_by_key": (C.c_int32, [c_vp, HazKeyC]),
    "jg_cde_save": (C.c_int32, [c_vp, C.POINTER(c_vp)]),
    "jg_cde_restore": (C.c_int32, [c_vp, c_vp]),
    "jg_snapshot_free": (None, [c_vp]),
    "jg_cde_bbox": (C.c_int32, [c_vp, c_f32p]),
    "jg_cde_hkey_exterior": (C.c_int32, [c_vp, C.POINTER(HazKeyC)]),
    "jg_cde_sync": (C.c_int32, [c_vp]),
    "jg_cde_get_stats": (C.c_int32, [c_vp, C.POINTER(CDEStatsC)]),
    "jg_cde_dump": (C.c_int32, [c_vp, c_u32p, C.c_uint64, c_u32p, C.c_uint64, c_u64p, c_u64p]),
    "jg_detect_batch": (C.c_int32, [c_vp, c_vp, c_vp, C.c_uint64, C.POINTER(FilterC), C.c_uint32, c_vp, c_vp]),
    "jg_collect_batch": (C.c_int32, [c_vp, c_vp, c_vp, C.c_uint64, C.c_uint32, c_vp, c_vp, c_vp, c_vp]),
    "jg_detect_batch_host": (C.c_int32, [c_vp, c_vp, c_f32p, C.c_uint64, C.POINTER(FilterC), C.c_uint32, c_u8p]),
    "jg_collect_batch_host": (C.c_int32, [c_vp, c_vp, c_f32p, C.c_uint64, C.c_uint32, c_u32p, c_u64p, c_u32p]),
    "jg_plan_create": (C.c_int32, [C.POINTER(QuerySetC), C.c_uint32, C.POINTER(PlanConfigC), C.POINTER(c_vp)]),
    "jg_plan_free": (None, [c_vp]),
    "jg_plan_detect": (C.c_int32, [c_vp, c_vp, c_vp, c_vp]),
    "jg_plan_detect_graph": (C.c_int32, [c_vp, c_vp, c_vp, C.c_uint32, c_vp]),
    "jg_plan_autotune": (C.c_int32, [c_vp, c_vp, c_vp, C.c_uint32, c_vp, C.POINTER(PlanConfigC),
                                     C.POINTER(C.c_float)]),
    "jg_plan_get_config": (C.c_int32, [c_vp, C.POINTER(PlanConfigC)]),
    "jg_plan_collect": (C.c_int32, [c_vp, c_vp, C.c_uint32, c_vp, c_vp, c_vp, c_vp]),
    "jg_detect_sets_noalloc": (C.c_int32, [C.POINTER(QuerySetC), C.c_uint32, C.POINTER(PlanConfigC), c_vp, c_vp,
                                           c_vp]),
    "jg_search_create": (C.c_int32, [C.POINTER(SearchItemC), C.c_uint32, C.POINTER(SearchConfigC), C.POINTER(c_vp)]),
    "jg_search_free": (None, [c_vp]),
    "jg_search_run": (C.c_int32, [c_vp, C.c_uint32, c_vp]),
    "jg_search_results": (C.c_int32, [c_vp, C.POINTER(SearchResultC)]),
    "jg_search_samples": (C.c_int32, [c_vp, c_f32p, c_u8p]),
    "jg_overlap_create": (C.c_int32, [c_f32p, c_u32p, C.c_uint32, c_u32p, C.c_uint64, c_f32p, C.c_uint32,
                                      C.POINTER(c_vp)]),
    "jg_overlap_set_pairs": (C.c_int32, [c_vp, c_u32p, C.c_uint64]),
    "jg_overlap_fwd": (C.c_int32, [c_vp, c_vp, c_vp, c_vp]),
    "jg_overlap_bwd": (C.c_int32, [c_vp, c_vp, c_vp, c_vp]),
    "jg_overlap_fwd_bwd": (C.c_int32, [c_vp, c_vp, c_vp, c_vp, c_vp]),
    "jg_overlap_free": (None, [c_vp]),
    "jg_overlap_pairs_from_cde": (C.c_int32, [c_vp, c_vp, c_u32p, c_vp, c_f32p, c_u64p, c_vp]),
    "jg_overlap_get_pairs": (C.c_int32, [c_vp, c_u32p, C.c_uint64, c_u64p]),
    "jg_overlap_candidate_pairs": (C.c_int32, [c_f32p, c_f32p, C.c_uint32, c_f32p, C.c_uint32, c_u32p, C.c_uint64,
                                               c_u64p]),
    "jg_graph_begin": (C.c_int32, [c_vp]),
    "jg_graph_end": (C.c_int32, [c_vp, C.POINTER(c_vp)]),
    "jg_graph_launch": (C.c_int32, [c_vp, c_vp]),
    "jg_graph_free": (None, [c_vp]),
    "jg_version": (C.c_char_p, []),
    "jg_last_error": (C.c_char_p, []),
}

EXPORTED = sorted(_SIGS)

_lib = None


def lib_path():
    """The in-tree build, or the library named by JAGUA_AMD_LIB (another build of the same ABI)."""
    return os.environ.get("JAGUA_AMD_LIB") or os.path.join(LIB_DIR, "libjagua_amd.so")


def lib():
    """The loaded libjagua_amd.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        p = lib_path()
        if not os.path.exists(p):
            raise RuntimeError(f"{p} is missing: build it with `python jagua-rs_amd/build.py` "
                               "(the collision engine has no CPU fallback)")
        L = C.CDLL(p)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class JaguaError(RuntimeError):
    pass


def check(status):
    if status != 0:
        msg = lib().jg_last_error().decode()
        raise JaguaError(f"jagua_amd error {status}: {msg}")


_synth = None


class LayoutParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("container_kind", C.c_uint32), ("cw", C.c_float), ("ch", C.c_float),
                ("container_nv_lo", C.c_uint32), ("container_nv_hi", C.c_uint32), ("n_holes", C.c_uint32),
                ("hole_kind", C.c_uint32), ("hole_lo", C.c_float), ("hole_hi", C.c_float),
                ("hole_nv_lo", C.c_uint32), ("hole_nv_hi", C.c_uint32), ("n_items", C.c_uint32),
                ("item_nv_lo", C.c_uint32), ("item_nv_hi", C.c_uint32), ("item_r_lo", C.c_float),
                ("item_r_hi", C.c_float), ("item_shape", C.c_uint32), ("min_sep", C.c_float),
                ("rot_kind", C.c_uint32), ("placement", C.c_uint32), ("max_attempts", C.c_uint32),
                ("lbf_samples", C.c_uint32)]


def synth():
    global _synth
    if _synth is None:
        p = os.path.join(LIB_DIR, "libjagua_synth.so")
        if not os.path.exists(p):
            raise RuntimeError(f"{p} is missing: build it with `python jagua-rs_amd/build.py`")
        L = C.CDLL(p)
        L.js_layout_generate.restype = c_vp
        L.js_layout_generate.argtypes = [C.POINTER(LayoutParams)]
        L.js_layout_free.argtypes = [c_vp]
        L.js_layout_sizes.argtypes = [c_vp, c_u32p]
        L.js_layout_container.argtypes = [c_vp, c_f32p]
        L.js_layout_hole_n.restype = C.c_uint32
        L.js_layout_hole_n.argtypes = [c_vp, C.c_uint32]
        L.js_layout_hole.argtypes = [c_vp, C.c_uint32, c_f32p]
        L.js_layout_item_n.restype = C.c_uint32
        L.js_layout_item_n.argtypes = [c_vp, C.c_uint32]
        L.js_layout_item.argtypes = [c_vp, C.c_uint32, c_f32p, c_f32p]
        L.js_polygon.restype = C.c_uint32
        L.js_polygon.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_uint32, c_f32p]
        L.js_transforms.argtypes = [C.c_uint64, C.c_uint64, C.c_float, C.c_float, C.c_float, C.c_float,
                                    C.c_uint32, c_f32p]
        _synth = L
    return _synth
