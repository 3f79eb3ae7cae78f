"""ctypes binding of libdiskann_mi355x.so (the C ABI in include/diskann_mi355x.h).

The library is loaded from this package directory; a missing library is an error (there
is no fallback path).
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdiskann_mi355x.so")
# hardware queues the serving lanes need: one per lane (16) plus the index and null streams
SERVE_HW_QUEUES = 20

# Every symbol declared in include/diskann_mi355x.h.
EXPORTS = [
    "dkm_last_error", "dkm_version", "dkm_init", "dkm_device_alloc", "dkm_device_free",
    "dkm_memcpy_h2d", "dkm_memcpy_d2h", "dkm_device_copy", "dkm_synchronize",
    "dkm_distance_batch",
    "dkm_pq_chunk_offsets", "dkm_pq_lut", "dkm_pq_encode", "dkm_pq_train",
    "dkm_pq_train_draws",
    "dkm_index_create", "dkm_index_destroy", "dkm_set_vectors", "dkm_set_start_points",
    "dkm_set_graph", "dkm_get_graph", "dkm_set_pq", "dkm_set_codes", "dkm_encode_index",
    "dkm_get_codes", "dkm_search_batch", "dkm_disk_search_batch", "dkm_search_batch_device", "dkm_index_synchronize",
    "dkm_profile", "dkm_profile_read", "dkm_set_tuning", "dkm_build", "dkm_prune_batch", "dkm_prune_range", "dkm_set_build_hybrid", "dkm_set_build_sq", "dkm_set_build_spherical", "dkm_medoid",
    "dkm_bruteforce", "dkm_closest_centers", "dkm_search_batch_device_stats", "dkm_autotune",
    "dkm_search_graph_create", "dkm_search_graph_launch", "dkm_search_graph_destroy",
    "dkm_search_status", "dkm_search_graph_status", "dkm_set_visited_capacity",
    "dkm_search_batch_device_noalloc",
    "dkm_server_start", "dkm_server_slots", "dkm_server_submit",
    "dkm_server_submit_batch", "dkm_server_poll", "dkm_server_pending", "dkm_server_stop",
    "dkm_index_shape", "dkm_bin_write", "dkm_bin_read_shape", "dkm_bin_read",
    "dkm_set_vectors_from_bin", "dkm_save_graph", "dkm_save_graph_disk", "dkm_load_graph",
    "dkm_pq_pivots_write", "dkm_pq_pivots_read_shape", "dkm_pq_pivots_read",
    "dkm_kmeans_device", "dkm_kmeans_assign_device", "dkm_pq_distance_batch",
    "dkm_prescore_create", "dkm_prescore_search", "dkm_prescore_destroy",
    "dkm_sq_train", "dkm_sq_compress", "dkm_sq_distance_batch", "dkm_set_sq",
    "dkm_encode_index_sq", "dkm_set_sq_codes", "dkm_get_sq_codes", "dkm_use_pq",
    "dkm_sph_create", "dkm_sph_destroy", "dkm_sph_dims", "dkm_sph_train_stats",
    "dkm_sph_compress", "dkm_sph_distance_batch", "dkm_sph_query_distance_batch",
    "dkm_set_spherical", "dkm_encode_index_sph", "dkm_set_sph_codes", "dkm_get_sph_codes",
    "dkm_minmax_create", "dkm_minmax_destroy", "dkm_minmax_dims", "dkm_minmax_compress",
    "dkm_minmax_distance_batch", "dkm_minmax_query_distance_batch",
]

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libdiskann_mi355x.so is not built; run `python -m diskann_amd.build` (no CPU "
            "fallback exists)")
    # the serving lanes (sixteen streams) run concurrently only with a hardware queue
    # each; HIP reads this once, when the runtime starts, and its default is 4 (streams
    # beyond the queues share them and their launches serialise)
    if int(os.environ.get("GPU_MAX_HW_QUEUES", "0") or 0) < SERVE_HW_QUEUES:
        os.environ["GPU_MAX_HW_QUEUES"] = str(SERVE_HW_QUEUES)
    lib = C.CDLL(LIB_PATH)
    _declare(lib)
    _lib = lib
    return lib


def _declare(L):
    P, I64, U32, I = C.c_void_p, C.c_int64, C.c_uint32, C.c_int
    P64, S = C.POINTER(C.c_int64), C.c_char_p
    L.dkm_last_error.restype = C.c_char_p
    L.dkm_version.restype = C.c_char_p
    sig = {
        "dkm_init": [I],
        "dkm_device_alloc": [C.c_size_t, C.POINTER(C.c_void_p)],
        "dkm_device_free": [P],
        "dkm_memcpy_h2d": [P, P, C.c_size_t],
        "dkm_memcpy_d2h": [P, P, C.c_size_t],
        "dkm_device_copy": [P, P, C.c_size_t],
        "dkm_synchronize": [],
        "dkm_distance_batch": [I, I64, P, I64, P, I64, P, I64, P],
        "dkm_pq_chunk_offsets": [I64, I64, P],
        "dkm_pq_lut": [I, P, I64, I64, P, I64, P, I64, P],
        "dkm_pq_encode": [P, I64, I64, P, I64, P, I64, P],
        "dkm_pq_train": [P, I64, I64, I64, I64, I64, C.c_uint64, P, P],
        "dkm_pq_train_draws": [P, I64, I64, I64, I64, I64, P, P, P],
        "dkm_pq_distance_batch": [I, P, I64, I64, P, I64, P, P, I64, P, I64, P, I64, P],
        "dkm_index_create": [I, I64, I64, I64, I64, I64, C.POINTER(C.c_void_p)],
        "dkm_set_vectors": [P, P, I64, I64],
        "dkm_set_start_points": [P, P],
        "dkm_set_graph": [P, P, P, I64, I64, I64],
        "dkm_get_graph": [P, P, P, I64, I64, I64],
        "dkm_set_pq": [P, P, I64, P, I64],
        "dkm_set_codes": [P, P, I64, I64],
        "dkm_encode_index": [P],
        "dkm_sq_train": [P, I64, I64, C.c_double, P, P, P],
        "dkm_sq_compress": [C.c_float, P, I64, P, I64, P],
        "dkm_sq_distance_batch": [I, C.c_float, P, I64, P, I64, P, I64, P, I64, P],
        "dkm_set_sq": [P, C.c_float, P, C.c_float],
        "dkm_encode_index_sq": [P],
        "dkm_set_sq_codes": [P, P, I64, I64],
        "dkm_get_sq_codes": [P, P, I64, I64],
        "dkm_use_pq": [P],
        "dkm_sph_create": [I, I64, P, C.c_float, C.c_float, I, P, P, I64, I64, P, I64,
                           C.POINTER(C.c_void_p)],
        "dkm_sph_dims": [P, C.POINTER(I64), C.POINTER(I64)],
        "dkm_sph_train_stats": [I, P, I64, I64, P, P],
        "dkm_sph_compress": [P, P, I64, P],
        "dkm_sph_distance_batch": [P, P, I64, P, I64, P, I64, P],
        "dkm_sph_query_distance_batch": [P, I, I, P, I64, P, I64, P, I64, P],
        "dkm_set_spherical": [P, P, I],
        "dkm_encode_index_sph": [P],
        "dkm_set_sph_codes": [P, P, I64, I64],
        "dkm_get_sph_codes": [P, P, I64, I64],
        "dkm_minmax_create": [I64, C.c_float, I, P, P, I64, I64, P, I64, C.POINTER(C.c_void_p)],
        "dkm_minmax_dims": [P, C.POINTER(I64), C.POINTER(I64)],
        "dkm_minmax_compress": [P, P, I64, P, P],
        "dkm_minmax_distance_batch": [P, I, P, I64, P, I64, P, I64, P],
        "dkm_minmax_query_distance_batch": [P, I, P, I64, P, I64, P, I64, P],
        "dkm_get_codes": [P, P, I64, I64],
        "dkm_search_batch": [P, I, P, I64, P, P, U32, I64, P, P, P, P, P],
        "dkm_disk_search_batch": [P, P, I64, P, P, U32, U32, C.c_uint64, I64, P, P, P, P, P],
        "dkm_search_batch_device": [P, I, P, I64, P, P, U32, U32, U32, I64, P, P, P, P, P],
        "dkm_index_synchronize": [P],
        "dkm_profile": [P, I],
        "dkm_profile_read": [P, C.POINTER(C.c_int64), C.POINTER(C.c_double)],
        "dkm_set_tuning": [P, I64, I],
        "dkm_build": [P, P, I64, I64, I64, U32, C.c_float, I64, I],
        "dkm_prune_batch": [P, P, I64, P, P, I64, C.c_float, I64, I, I, P, P, I64],
        "dkm_prune_range": [P, P, I64, C.c_float, I64, I],
        "dkm_set_build_hybrid": [P, I64],
        "dkm_set_build_sq": [P, I],
        "dkm_set_build_spherical": [P, I],
        "dkm_medoid": [P, P, C.POINTER(C.c_int64)],
        "dkm_bruteforce": [P, P, I64, I64, P, P],
        "dkm_closest_centers": [P, I64, I64, P, I64, I64, P],
        "dkm_autotune": [P, I, P, I64, U32, U32, U32, C.POINTER(C.c_int64), C.POINTER(C.c_int)],
        "dkm_server_start": [P, I, U32, U32, U32, I64, I64, C.POINTER(C.c_void_p)],
        "dkm_server_slots": [P, C.POINTER(C.c_int64)],
        "dkm_server_submit": [P, P, U32, U32, C.c_uint64],
        "dkm_server_submit_batch": [P, P, I64, P, P, P, C.POINTER(C.c_int64)],
        "dkm_server_poll": [P, I64, P, P, P, P, P, P, P, C.POINTER(C.c_int64)],
        "dkm_server_pending": [P, C.POINTER(C.c_int64)],
        "dkm_server_stop": [P],
        "dkm_index_shape": [P, P64, P64, P64, P64, P64],
        "dkm_bin_write": [S, P, I64, I64, I64],
        "dkm_bin_read_shape": [S, P64, P64],
        "dkm_bin_read": [S, P, I64, I64, I64],
        "dkm_set_vectors_from_bin": [P, S],
        "dkm_save_graph": [P, S, U32],
        "dkm_save_graph_disk": [P, S, U32],
        "dkm_load_graph": [P, S, C.POINTER(C.c_uint32)],
        "dkm_pq_pivots_write": [S, P, I64, I64, P, P, I64],
        "dkm_pq_pivots_read_shape": [S, P64, P64, P64],
        "dkm_pq_pivots_read": [S, P, I64, I64, P, I64],
        "dkm_kmeans_device": [P, I64, I64, I64, I64, I, P, P, C.POINTER(C.c_double),
                              C.POINTER(C.c_double)],
        "dkm_kmeans_assign_device": [P, I64, I64, P, I64, I, P, P],
        "dkm_prescore_create": [P, I64, I64, I, I, C.POINTER(C.c_void_p)],
        "dkm_prescore_search": [P, P, I64, I64, I64, P, P, C.POINTER(C.c_double)],
        "dkm_search_batch_device_stats": [P, I, P, I64, P, P, U32, U32, U32, I64, P, P, P, P, P,
                                          P],
        "dkm_search_graph_create": [P, I, P, I64, P, P, U32, U32, U32, I64, P, P, P, P, P, P,
                                    C.POINTER(C.c_void_p)],
        "dkm_search_graph_launch": [P],
        "dkm_search_graph_status": [P, P],
        "dkm_search_status": [P, P],
        "dkm_search_batch_device_noalloc": [P, I, P, I64, P, P, U32, U32, U32, I64, P, P, P, P,
                                            P, P],
        "dkm_set_visited_capacity": [P, I64],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.dkm_index_destroy.argtypes = [P]
    L.dkm_index_destroy.restype = None
    L.dkm_search_graph_destroy.argtypes = [P]
    L.dkm_search_graph_destroy.restype = None
    L.dkm_prescore_destroy.argtypes = [P]
    L.dkm_prescore_destroy.restype = None
    L.dkm_sph_destroy.argtypes = [P]
    L.dkm_sph_destroy.restype = None
    L.dkm_minmax_destroy.argtypes = [P]
    L.dkm_minmax_destroy.restype = None
