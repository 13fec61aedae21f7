"""ctypes binding of libfmmhip.so (the C ABI declared in include/fmmhip.h).

The library is built in-tree (fastmultipole_amd/libfmmhip.so) by
__graft_entry__.build() or `make -C fastmultipole_amd/csrc`. There is no CPU
fallback: every call below fails loudly if the library or the GPU is missing.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfmmhip.so")

FMMHIP_OK, FMMHIP_EINVAL, FMMHIP_ENOMEM, FMMHIP_EDEVICE, FMMHIP_ETOL, FMMHIP_EDEPTH = range(6)
POINT_SOURCE, POINT_VORTEX, POINT_DIPOLE, VORTEX_GAUSSIAN = 0, 1, 2, 3
PANEL_SOURCE_TRI, PANEL_DIPOLE_TRI, PANEL_SOURCE_QUAD, PANEL_DIPOLE_QUAD = 4, 5, 6, 7
POINT_SOURCE_VORTEX = 8
PANEL_SOURCE_DIPOLE_TRI, PANEL_SOURCE_DIPOLE_QUAD = 9, 10


def data_per_body(kernel):
    """Rows per body of a kernel's source buffer: [x, y, z, radius, strength rows...]; panels add their
    vertices (3 or 4 x 3 rows)."""
    if kernel == POINT_SOURCE_VORTEX:
        return 8
    if kernel in (PANEL_SOURCE_DIPOLE_TRI, PANEL_SOURCE_DIPOLE_QUAD):
        return 6 + 3 * (4 if kernel == PANEL_SOURCE_DIPOLE_QUAD else 3)
    if kernel >= PANEL_SOURCE_TRI:
        return 5 + 3 * (4 if kernel >= PANEL_SOURCE_QUAD else 3)
    return 5 if kernel == POINT_SOURCE else 7


FP64, MIXED = 0, 1
INFLUENCE_POTENTIAL, INFLUENCE_NORMAL_GRADIENT = 0, 1
BARBA, SELFTUNING, SELFTUNING_TREESTOP = 0, 1, 2
ERR_NONE, ERR_ROTATED_COEFFICIENTS_GRADIENT, ERR_POWER_POTENTIAL, ERR_POWER_GRADIENT = 0, 1, 2, 3

EXPORTED = [
    "fmmhip_options_default", "fmmhip_ctx_create", "fmmhip_ctx_destroy", "fmmhip_last_error", "fmmhip_sync",
    "fmmhip_ctx_set_p2p_free_slots",
    "fmmhip_eval_create", "fmmhip_eval_destroy", "fmmhip_eval_set_systems", "fmmhip_eval_set_batch",
    "fmmhip_eval_get_batch_stats", "fmmhip_eval_set_tuple", "fmmhip_eval_set_tuple_leaf_sizes", "fmmhip_eval_build_trees",
    "fmmhip_eval_build_lists", "fmmhip_eval_run", "fmmhip_fmm", "fmmhip_eval_get_targets", "fmmhip_eval_n_branches",
    "fmmhip_eval_n_levels", "fmmhip_eval_n_leaves", "fmmhip_eval_get_branches", "fmmhip_eval_get_levels",
    "fmmhip_eval_get_leaf_index", "fmmhip_eval_get_sort_index", "fmmhip_eval_n_list", "fmmhip_eval_get_list",
    "fmmhip_eval_get_m2l_orders", "fmmhip_eval_get_expansions", "fmmhip_direct", "fmmhip_direct_noalloc",
    "fmmhip_fgs_setup", "fmmhip_fgs_solve", "fmmhip_fgs_n_leaves", "fmmhip_fgs_leaf_info", "fmmhip_fgs_get_matrix",
    "fmmhip_fgs_nonself_rows",
    "fmmhip_server_create", "fmmhip_server_destroy", "fmmhip_server_set_batching", "fmmhip_server_submit", "fmmhip_server_wait",
    "fmmhip_server_stats",
]
REQUEST_FMM, REQUEST_SOLVE = 0, 1


class Options(C.Structure):
    _fields_ = [
        ("expansion_order", C.c_int32), ("leaf_size_source", C.c_int32), ("leaf_size_target", C.c_int32),
        ("multipole_acceptance", C.c_double), ("list_method", C.c_int32), ("shrink_recenter", C.c_int32),
        ("error_kind", C.c_int32), ("error_relative", C.c_int32), ("error_tolerance", C.c_double),
        ("error_bonus", C.c_int32), ("scalar_potential", C.c_int32), ("gradient", C.c_int32), ("hessian", C.c_int32),
        ("farfield", C.c_int32), ("nearfield", C.c_int32), ("self_induced", C.c_int32), ("precision", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("pmax", C.c_int32), ("error_success", C.c_int32), ("n_target_branches", C.c_int64),
        ("n_source_branches", C.c_int64), ("n_m2l", C.c_int64), ("n_direct", C.c_int64), ("n_p2p_pairs", C.c_int64),
        ("ms_tree", C.c_double), ("ms_lists", C.c_double), ("ms_p2p", C.c_double), ("ms_upward", C.c_double),
        ("ms_m2l", C.c_double), ("ms_downward", C.c_double), ("ms_total", C.c_double),
        ("ms_estimate", C.c_double), ("ms_gather", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Request(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("kernel", C.c_int32), ("n", C.c_int64), ("source_buffer", C.c_void_p),
        ("options", Options), ("influence", C.c_int32), ("max_iterations", C.c_int32), ("tolerance", C.c_double),
        ("external", C.c_void_p),
    ]


class RequestStats(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("sweeps", C.c_int32), ("pmax", C.c_int32), ("error_success", C.c_int32),
        ("converged", C.c_int32), ("queue_ms", C.c_double),
        ("run_ms", C.c_double), ("latency_ms", C.c_double), ("start_ms", C.c_double), ("end_ms", C.c_double),
        ("footprint_bytes", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerStats(C.Structure):
    _fields_ = [
        ("completed", C.c_int64), ("queued", C.c_int64), ("running", C.c_int32), ("p50_ms", C.c_double),
        ("p99_ms", C.c_double), ("mean_queue_ms", C.c_double), ("span_ms", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FmmHipError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"libfmmhip status {status}: {msg}")
        self.status = status


_LIB = None


def load():
    """Load libfmmhip.so (no compute call; usable without a GPU)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "or `make -C fastmultipole_amd/csrc` (there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dp = C.c_void_p, C.c_int32, C.c_int64, C.c_void_p
    L.fmmhip_options_default.argtypes = [C.POINTER(Options)]
    L.fmmhip_options_default.restype = None
    L.fmmhip_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.fmmhip_ctx_destroy.argtypes = [vp]
    L.fmmhip_ctx_destroy.restype = None
    L.fmmhip_last_error.argtypes = [vp]
    L.fmmhip_last_error.restype = C.c_char_p
    L.fmmhip_sync.argtypes = [vp]
    L.fmmhip_ctx_set_p2p_free_slots.argtypes = [vp, i32]
    L.fmmhip_eval_create.argtypes = [vp, C.POINTER(vp)]
    L.fmmhip_eval_destroy.argtypes = [vp]
    L.fmmhip_eval_destroy.restype = None
    L.fmmhip_eval_set_systems.argtypes = [vp, i32, i64, dp, i64, dp, i32]
    L.fmmhip_eval_set_batch.argtypes = [vp, i32, i64, dp, dp, i32]
    L.fmmhip_eval_get_batch_stats.argtypes = [vp, dp, dp]
    L.fmmhip_eval_set_tuple.argtypes = [vp, i64, dp, dp, dp, i32]
    L.fmmhip_eval_set_tuple_leaf_sizes.argtypes = [vp, i64, dp]
    for f in ("fmmhip_eval_build_trees", "fmmhip_eval_build_lists"):
        getattr(L, f).argtypes = [vp, C.POINTER(Options)]
    for f in ("fmmhip_eval_run", "fmmhip_fmm"):
        getattr(L, f).argtypes = [vp, C.POINTER(Options), C.POINTER(Stats)]
    L.fmmhip_eval_get_targets.argtypes = [vp, dp, i32]
    for f in ("fmmhip_eval_n_branches", "fmmhip_eval_n_levels", "fmmhip_eval_n_leaves", "fmmhip_eval_n_list"):
        getattr(L, f).argtypes = [vp, i32]
        getattr(L, f).restype = i64
    L.fmmhip_eval_get_branches.argtypes = [vp, i32, dp, dp]
    L.fmmhip_eval_get_levels.argtypes = [vp, i32, dp]
    L.fmmhip_eval_get_leaf_index.argtypes = [vp, i32, dp]
    L.fmmhip_eval_get_sort_index.argtypes = [vp, i32, dp]
    L.fmmhip_eval_get_list.argtypes = [vp, i32, dp]
    L.fmmhip_eval_get_m2l_orders.argtypes = [vp, dp]
    L.fmmhip_eval_get_expansions.argtypes = [vp, i32, dp]
    L.fmmhip_direct.argtypes = [vp, i32, i64, dp, i64, dp, i32, i32, i32, i32, dp, i32]
    L.fmmhip_direct_noalloc.argtypes = [vp, i32, i64, dp, i64, dp, i32, i32, i32, i32, dp]
    L.fmmhip_fgs_setup.argtypes = [vp, C.POINTER(Options), i32]
    L.fmmhip_fgs_solve.argtypes = [vp, dp, i32, i32, C.c_double, dp, C.POINTER(i32), dp]
    L.fmmhip_fgs_n_leaves.argtypes = [vp]
    L.fmmhip_fgs_n_leaves.restype = i64
    L.fmmhip_fgs_leaf_info.argtypes = [vp, dp]
    L.fmmhip_fgs_get_matrix.argtypes = [vp, i32, i64, dp]
    L.fmmhip_fgs_nonself_rows.argtypes = [vp, i64, dp]
    L.fmmhip_server_create.argtypes = [C.c_int, i32, C.c_double, C.POINTER(vp)]
    L.fmmhip_server_destroy.argtypes = [vp]
    L.fmmhip_server_destroy.restype = None
    L.fmmhip_server_set_batching.argtypes = [vp, i32, i64]
    L.fmmhip_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(i64)]
    L.fmmhip_server_wait.argtypes = [vp, i64, dp, dp, C.POINTER(RequestStats)]
    L.fmmhip_server_stats.argtypes = [vp, C.POINTER(ServerStats)]
    _LIB = L
    return L


def default_options():
    o = Options()
    load().fmmhip_options_default(C.byref(o))
    return o
