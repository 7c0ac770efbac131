"""ctypes binding of libfclx.so (the C ABI declared in include/fclx.h and
include/fclx_synth.h).

The library is loaded from this package directory.  There is no fallback: if
it is missing the import of any entry point raises NativeLibraryMissing.
"""
import ctypes
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfclx.so")

FCLX_OK = 0
FCLX_ERR_INVALID_ARGUMENT = -1
FCLX_ERR_OUT_OF_MEMORY = -2
FCLX_ERR_CAPACITY = -3
FCLX_ERR_DEVICE = -4
FCLX_ERR_UNSUPPORTED = -5
FCLX_ERR_EMPTY_MODEL = -6
FCLX_ERR_STACK_OVERFLOW = -7

FCLX_BV_OBBRSS = 0
FCLX_BV_RSS = 1
FCLX_SHAPE_BOX = 0
FCLX_SHAPE_SPHERE = 1
FCLX_SHAPE_CAPSULE = 2
FCLX_SHAPE_CYLINDER = 3
FCLX_SHAPE_CONVEX = 4
FCLX_ENGINE_PLAIN = 0
FCLX_ENGINE_FAST = 1
FCLX_ENGINE_BOOL = 2

c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_u64 = ctypes.c_uint64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
P_i32 = ctypes.POINTER(c_i32)
P_i64 = ctypes.POINTER(c_i64)
P_dbl = ctypes.POINTER(c_dbl)


class NativeLibraryMissing(RuntimeError):
    pass


class MeshInfo(ctypes.Structure):
    _fields_ = [("num_bvs", c_i32), ("num_tris", c_i32), ("num_vertices", c_i32),
                ("depth", c_i32), ("bv_type", c_i32), ("reserved", c_i32),
                ("device_bytes", c_i64)]


class CollideBatchArgs(ctypes.Structure):
    _fields_ = [("num_queries", c_i32), ("query_meshes", c_vp), ("tf1", c_vp), ("tf2", c_vp),
                ("num_max_contacts", c_i32), ("engine", c_i32), ("counts", c_vp),
                ("pairs", c_vp), ("pair_capacity", c_i64), ("num_pairs", c_vp),
                ("bv_tests", c_vp), ("leaf_tests", c_vp), ("enable_contact", c_i32),
                ("reserved", c_i32), ("contact_pos", c_vp), ("contact_normal", c_vp),
                ("contact_depth", c_vp)]


class DistanceBatchArgs(ctypes.Structure):
    _fields_ = [("num_queries", c_i32), ("query_meshes", c_vp), ("tf1", c_vp), ("tf2", c_vp),
                ("enable_nearest_points", c_i32), ("engine", c_i32), ("distance", c_vp),
                ("points", c_vp), ("ids", c_vp), ("bv_tests", c_vp), ("leaf_tests", c_vp)]


class ShapeRec(ctypes.Structure):
    _fields_ = [("type", c_i32), ("num_vertices", c_i32), ("dims", c_dbl * 3),
                ("vertex_offset", c_i64)]


class ShapeCollideBatchArgs(ctypes.Structure):
    _fields_ = [("num_pairs", c_i32), ("num_shapes", c_i32), ("shapes", c_vp),
                ("vertices", c_vp), ("num_vertices", c_i64), ("pair_shapes", c_vp),
                ("tf1", c_vp), ("tf2", c_vp), ("num_max_contacts", c_i32),
                ("enable_contact", c_i32), ("gjk_tolerance", c_dbl), ("counts", c_vp),
                ("contact_pos", c_vp), ("contact_normal", c_vp), ("contact_depth", c_vp),
                ("passes", c_vp)]


class MeshShapeCollideBatchArgs(ctypes.Structure):
    _fields_ = [("num_queries", c_i32), ("query_meshes", c_vp), ("query_shapes", c_vp),
                ("num_shapes", c_i32), ("shapes", c_vp), ("vertices", c_vp),
                ("num_vertices", c_i64), ("tf1", c_vp), ("tf2", c_vp),
                ("num_max_contacts", c_i32), ("enable_contact", c_i32),
                ("gjk_tolerance", c_dbl), ("counts", c_vp), ("primitives", c_vp),
                ("contact_pos", c_vp), ("contact_normal", c_vp), ("contact_depth", c_vp),
                ("stats", c_vp)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("scene_mesh", c_i32), ("num_streams", c_i32), ("scene_pose", c_dbl * 12),
                ("hbm_budget", c_i64), ("max_batch_queries", c_i32), ("engine", c_i32)]


class Request(ctypes.Structure):
    _fields_ = [("num_links", c_i32), ("num_waypoints", c_i32), ("link_meshes", c_vp),
                ("link_poses", c_vp), ("result", c_vp), ("tag", c_i64)]


class Completion(ctypes.Structure):
    _fields_ = [("tag", c_i64), ("status", c_i32), ("num_queries", c_i32), ("queue_ms", c_dbl),
                ("latency_ms", c_dbl)]


class EngineReport(ctypes.Structure):
    _fields_ = [("setup_ms", c_dbl), ("expand_ms", c_dbl), ("leaves_ms", c_dbl),
                ("dfs_ms", c_dbl), ("fixup_ms", c_dbl), ("expand_bv_tests", c_i64),
                ("expand_leaf_tests", c_i64), ("dfs_bv_tests", c_i64), ("dfs_leaf_tests", c_i64),
                ("fp64_fallbacks", c_i64), ("front_size", c_i64), ("levels", c_i32),
                ("flagged", c_i32), ("dfs_waves", c_i32), ("dfs_rounds", c_i32),
                ("dfs_wave_busy", c_dbl), ("dfs_end_p50_ms", c_dbl), ("dfs_end_p90_ms", c_dbl),
                ("dfs_end_max_ms", c_dbl), ("bool_steps", c_i64 * 3), ("bool_lanes", c_i64 * 3),
                ("bool_skipped", c_i64), ("bool_rounds", c_i64), ("bool_stack_max", c_i64),
                ("bool_wave_ms", c_dbl), ("bool_end_p50_ms", c_dbl), ("bool_end_p90_ms", c_dbl),
                ("bool_span_ms", c_dbl), ("dfs_inline_leaf_tests", c_i64),
                ("dfs_wave_steps", c_i64)]


# name -> (restype, argtypes)
class FrontBatchArgs(ctypes.Structure):
    _fields_ = [("num_sequences", c_i32), ("seq_offsets", c_vp), ("query_meshes", c_vp),
                ("tf1", c_vp), ("tf2", c_vp), ("num_max_contacts", c_i32), ("counts", c_vp),
                ("pairs", c_vp), ("pair_capacity", c_i64), ("num_pairs", c_vp),
                ("bv_tests", c_vp), ("leaf_tests", c_vp)]


SIGNATURES = {
    "fclx_set_profiling": (c_i32, [c_vp, c_i32]),
    "fclx_engine_last_report": (c_i32, [c_vp, ctypes.POINTER(EngineReport)]),
    "fclx_set_engine_param": (c_i32, [c_vp, ctypes.c_char_p, c_dbl]),
    "fclx_get_engine_param": (c_i32, [c_vp, ctypes.c_char_p, ctypes.POINTER(c_dbl)]),
    "fclx_init": (c_i32, [c_i32, ctypes.POINTER(c_vp)]),
    "fclx_shutdown": (None, [c_vp]),
    "fclx_last_error": (ctypes.c_char_p, [c_vp]),
    "fclx_mesh_create": (c_i32, [c_vp, c_i32, P_dbl, c_i32, P_i32, c_i32, P_i32]),
    "fclx_mesh_destroy": (c_i32, [c_vp, c_i32]),
    "fclx_mesh_get_info": (c_i32, [c_vp, c_i32, ctypes.POINTER(MeshInfo)]),
    "fclx_mesh_export_nodes": (c_i32, [c_vp, c_i32, P_dbl, P_i32, P_dbl]),
    "fclx_bvh_build": (c_i32, [c_i32, P_dbl, c_i32, P_i32, c_i32, P_dbl, P_i32, P_dbl,
                               P_i32]),
    "fclx_collide_batch": (c_i32, [c_vp, ctypes.POINTER(CollideBatchArgs), c_vp]),
    "fclx_collide_batch_noalloc": (c_i32, [c_vp, ctypes.POINTER(CollideBatchArgs), c_vp]),
    "fclx_distance_batch": (c_i32, [c_vp, ctypes.POINTER(DistanceBatchArgs), c_vp]),
    "fclx_sync": (c_i32, [c_vp, c_vp]),
    "fclx_shape_collide_batch": (c_i32, [c_vp, ctypes.POINTER(ShapeCollideBatchArgs), c_vp]),
    "fclx_mesh_shape_collide_batch": (c_i32, [c_vp, ctypes.POINTER(MeshShapeCollideBatchArgs),
                                              c_vp]),
    "fclx_shape_collide": (c_i32, [c_vp, ctypes.POINTER(ShapeRec), P_dbl, P_dbl,
                                   ctypes.POINTER(ShapeRec), P_dbl, P_dbl, c_i32, c_i32, c_dbl,
                                   P_dbl, P_dbl, P_dbl, P_i32, P_i32]),
    "fclx_collide": (c_i32, [c_vp, c_i32, P_dbl, c_i32, P_dbl, c_i32, P_i32, c_i32, P_i32,
                              P_i64]),
    "fclx_distance": (c_i32, [c_vp, c_i32, P_dbl, c_i32, P_dbl, c_i32, P_dbl, P_dbl, P_i32,
                              P_i64]),
    "fclx_collide_contact": (c_i32, [c_vp, c_i32, P_dbl, c_i32, P_dbl, c_i32, P_i32, P_dbl,
                                     P_dbl, P_dbl, c_i32, P_i32, P_i64]),
    "fclx_server_create": (c_i32, [c_vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(c_vp)]),
    "fclx_server_submit": (c_i32, [c_vp, ctypes.POINTER(Request)]),
    "fclx_server_poll": (c_i32, [c_vp, ctypes.POINTER(Completion), c_i32, c_i32, P_i32]),
    "fclx_server_replay": (c_i32, [c_vp, ctypes.POINTER(Request), P_dbl, c_i32, P_dbl, P_dbl,
                                   P_dbl]),
    "fclx_server_stats": (c_i32, [c_vp, P_i64, P_i64]),
    "fclx_server_destroy": (c_i32, [c_vp]),
    "fclx_mesh_refit": (c_i32, [c_vp, c_i32, c_vp, c_i32, c_vp]),
    "fclx_broadphase_pairs": (c_i32, [c_vp, c_i32, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp]),
    "fclx_front_set_create": (c_i32, [c_vp, c_i32, c_i32, ctypes.POINTER(c_vp)]),
    "fclx_front_set_destroy": (c_i32, [c_vp]),
    "fclx_front_set_clear": (c_i32, [c_vp, c_vp]),
    "fclx_front_set_read": (c_i32, [c_vp, c_i32, P_i32, c_i32, P_i32]),
    "fclx_collide_front_batch": (c_i32, [c_vp, ctypes.POINTER(FrontBatchArgs), c_vp, c_vp]),
    "fclx_synth_icosphere_size": (None, [c_i32, P_i32, P_i32]),
    "fclx_synth_icosphere": (c_i32, [c_i32, c_dbl, c_u64, P_dbl, P_i32]),
    "fclx_synth_heightfield_size": (None, [c_i32, P_i32, P_i32]),
    "fclx_synth_heightfield": (c_i32, [c_i32, c_dbl, c_u64, P_dbl, P_i32]),
    "fclx_synth_scale_to_radius": (None, [P_dbl, c_i32, c_dbl]),
    "fclx_synth_poses": (c_i32, [c_i32, c_i32, c_dbl, c_dbl, c_u64, P_dbl, P_dbl]),
    "fclx_synth_uniform": (None, [c_i64, c_u64, P_dbl]),
}

_lib = None


def _one_hip_runtime():
    """torch ships its own HIP runtime with the same soname as the system's
    (libamdhip64.so.7) but links it by file name.  Loading torch first makes
    libfclx.so bind to that already-loaded runtime, so the process holds one
    HIP runtime and torch streams/pointers are valid in both."""
    if os.environ.get("FCLX_NO_TORCH") == "1":
        return
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """The loaded libfclx.so with argtypes set; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                "libfclx.so not found at %s; build it with __graft_entry__.build()" % LIB_PATH)
        _one_hip_runtime()
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def ptr(arr, ctype):
    """ctypes pointer to a C-contiguous numpy array."""
    assert arr.flags["C_CONTIGUOUS"]
    return arr.ctypes.data_as(ctypes.POINTER(ctype))
