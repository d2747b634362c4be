"""ctypes binding of libgpumlib_nmf.so (the C-ABI in include/gpumlib_nmf.h and
include/gpumlib_reduction.h).

The library is built in-tree by gpumlib_amd/build.py.  Loading fails loudly when
it is missing: there is no fallback implementation.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libgpumlib_nmf.so")

EUCLIDEAN, KL = 0, 1
ADDITIVE_EUCLIDEAN, ADDITIVE_KL, ADDITIVE_EUCLIDEAN_WITH_H = 2, 3, 4
FP32, BF16, FP8, FP32X = 0, 1, 2, 3
BATCH_SPLIT_K, BATCH_FUSED = 0, 1
OBJ_NONE, OBJ_FROBENIUS, OBJ_KL, OBJ_QUALITY = 0, 1, 2, 3
STREAM_AUTO, STREAM_NONTEMPORAL, STREAM_CACHED = 0, 1, 2
REFERENCE_EPS = 1e-8  # float(10e-9f) == float(1e-8f)

_c = ctypes
_vp, _i, _l, _f, _d, _sz, _u64 = _c.c_void_p, _c.c_int, _c.c_long, _c.c_float, _c.c_double, _c.c_size_t, _c.c_uint64


class Dim3(ctypes.Structure):
    _fields_ = [("x", _c.c_uint32), ("y", _c.c_uint32), ("z", _c.c_uint32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", _c.c_size_t), ("max_active", _c.c_int), ("round_iterations", _c.c_int),
                ("use_graph", _c.c_int), ("target_workgroups", _c.c_int), ("autotune", _c.c_int),
                ("split_k", _c.c_int), ("sequential", _c.c_int)]


class JobResult(ctypes.Structure):
    _fields_ = [("iterations", _c.c_int), ("converged", _c.c_int), ("objective", _c.c_double),
                ("submit_s", _c.c_double), ("admit_s", _c.c_double), ("finish_s", _c.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", _c.c_longlong), ("completed", _c.c_longlong), ("rounds", _c.c_longlong),
                ("iterations_batched", _c.c_longlong), ("peak_active", _c.c_int), ("pool_bytes", _c.c_size_t),
                ("pool_high_water", _c.c_size_t), ("autotunes", _c.c_longlong), ("tuned_target", _c.c_int),
                ("tuned_pass_grid", _c.c_int), ("tuned_update_grid", _c.c_int), ("tuned_engine", _c.c_int),
                ("tuned_split_ranks", _c.c_int), ("tuned_bucket_targets", _c.c_int * 3),
                ("tuned_bucket_tile_cols", _c.c_int * 3), ("tuned_bucket_split_ranks", _c.c_int * 3)]


class BatchTuning(ctypes.Structure):
    _fields_ = [("engine", _c.c_int), ("target_workgroups", _c.c_int), ("pass_grid", _c.c_int), ("update_grid", _c.c_int),
                ("ms_per_iteration", _c.c_double), ("ms_per_iteration_before", _c.c_double), ("trials", _c.c_int),
                ("cached", _c.c_int), ("split_ranks", _c.c_int), ("bucket_targets", _c.c_int * 3),
                ("bucket_tile_cols", _c.c_int * 3), ("bucket_split_ranks", _c.c_int * 3)]


# name -> (restype, argtypes)
SIGNATURES = {
    "gpumlib_cuda_nmf_multiplicative_euclidean_update_matrix": (_i, [_c.c_uint, _c.c_uint, _vp, _vp, _vp, _i]),
    "gpumlib_cuda_nmf_multiplicative_divergence_update_matrix_w": (_i, [Dim3, Dim3, _vp, _vp, _vp, _vp, _vp, _i, _i, _i]),
    "gpumlib_cuda_nmf_multiplicative_divergence_update_matrix_h": (_i, [Dim3, Dim3, _vp, _vp, _vp, _vp, _vp, _i, _i, _i]),
    "gpumlib_cuda_nmf_additive_euclidean_update_matrix": (_i, [_c.c_uint, _c.c_uint, _vp, _vp, _vp, _i]),
    "gpumlib_cuda_nmf_additive_divergence_update_matrix_h": (_i, [Dim3, Dim3, _vp, _vp, _vp, _vp, _vp, _i, _i, _i]),
    "gpumlib_cuda_nmf_additive_divergence_update_matrix_w": (_i, [Dim3, Dim3, _vp, _vp, _vp, _vp, _vp, _i, _i, _i]),
    "gpumlib_nmf_additive_update": (_i, [_vp, _vp, _vp, _l, _vp]),
    "gpumlib_nmf_additive_kl_update_h": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _f, _vp]),
    "gpumlib_nmf_additive_kl_update_w": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _f, _vp]),
    "gpumlib_nmf_mu_update": (_i, [_vp, _vp, _vp, _l, _f, _vp]),
    "gpumlib_nmf_sum_w": (_i, [_vp, _i, _i, _f, _vp, _vp]),
    "gpumlib_nmf_sum_h": (_i, [_vp, _i, _i, _f, _vp, _vp]),
    "gpumlib_nmf_kl_update_h": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _f, _vp]),
    "gpumlib_nmf_kl_update_w": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _i, _f, _vp]),
    "gpumlib_nmf_quality_workspace_bytes": (_sz, [_l]),
    "gpumlib_nmf_quality": (_i, [_vp, _vp, _l, _f, _vp, _sz, _vp, _vp]),
    "gpumlib_nmf_quality_f32": (_i, [_vp, _vp, _l, _f, _vp, _sz, _vp, _vp]),
    "gpumlib_nmf_sgemm": (_i, [_c.c_char, _c.c_char, _i, _i, _i, _f, _vp, _i, _vp, _i, _f, _vp, _i, _vp]),
    "gpumlib_nmf_gemm_bf16": (_i, [_c.c_char, _c.c_char, _i, _i, _i, _f, _vp, _i, _vp, _i, _f, _vp, _i, _vp]),
    "gpumlib_nmf_gemm_fp8": (_i, [_c.c_char, _c.c_char, _i, _i, _i, _f, _vp, _i, _i, _vp, _i, _i, _f, _vp, _i, _vp]),
    "gpumlib_nmf_set_stream_policy": (_i, [_i]),
    "gpumlib_nmf_workspace_bytes": (_sz, [_i, _i, _i, _i, _i]),
    "gpumlib_nmf_vstore_bytes": (_sz, [_i, _i, _i]),
    "gpumlib_nmf_vstore_pack": (_i, [_i, _i, _i, _vp, _vp, _sz, _vp]),
    "gpumlib_nmf_iterate": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _f, _i, _i, _i, _vp, _vp, _sz, _vp]),
    "gpumlib_nmf_iterate_noalloc": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _f, _i, _i, _i, _vp, _vp]),
    "gpumlib_nmf_objective": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _f, _i, _vp, _vp, _sz, _vp]),
    "gpumlib_nmf_pass": (_i, [_i, _i, _i, _i, _i, _i, _vp, _vp, _vp, _f, _vp, _sz, _vp]),
    "gpumlib_nmf_object_create": (_i, [_i, _i, _i, _i, _i, _vp, _vp, _vp, _c.POINTER(_vp)]),
    "gpumlib_nmf_object_set_epsilon": (_i, [_vp, _f]),
    "gpumlib_nmf_object_set_quality_interval": (_i, [_vp, _i]),
    "gpumlib_nmf_object_do_iteration": (_i, [_vp, _i]),
    "gpumlib_nmf_object_do_iterations": (_i, [_vp, _i, _i]),
    "gpumlib_nmf_object_quality_improvement": (_f, [_vp]),
    "gpumlib_nmf_object_get_w": (_i, [_vp, _vp]),
    "gpumlib_nmf_object_get_h": (_i, [_vp, _vp]),
    "gpumlib_nmf_object_get_wh": (_i, [_vp, _vp]),
    "gpumlib_nmf_object_objective": (_i, [_vp, _i, _c.POINTER(_d)]),
    "gpumlib_nmf_object_destroy": (None, [_vp]),
    "gpumlib_nmf_batch_create": (_i, [_i, _c.POINTER(_vp)]),
    "gpumlib_nmf_batch_destroy": (None, [_vp]),
    "gpumlib_nmf_batch_set_target": (_i, [_vp, _i]),
    "gpumlib_nmf_batch_set_engine": (_i, [_vp, _i]),
    "gpumlib_nmf_batch_engine": (_i, [_vp]),
    "gpumlib_nmf_batch_set_rank_split": (_i, [_vp, _i]),
    "gpumlib_nmf_batch_rank_split": (_i, [_vp]),
    "gpumlib_nmf_batch_set_tile_cols": (_i, [_vp, _i]),
    "gpumlib_nmf_batch_workspace_bytes": (_sz, [_vp, _i, _i, _i, _i]),
    "gpumlib_nmf_batch_add": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _f, _vp, _sz, _c.POINTER(_i)]),
    "gpumlib_nmf_batch_remove": (_i, [_vp, _i]),
    "gpumlib_nmf_batch_active": (_i, [_vp]),
    "gpumlib_nmf_batch_iterate": (_i, [_vp, _i, _i, _vp]),
    "gpumlib_nmf_batch_graph_launch": (_i, [_vp, _i, _i, _vp]),
    "gpumlib_nmf_batch_objectives": (_i, [_vp, _c.POINTER(_d), _vp]),
    "gpumlib_nmf_batch_objective_device_ptr": (_vp, [_vp]),
    "gpumlib_nmf_batch_autotune": (_i, [_vp, _i, _vp, _c.POINTER(BatchTuning)]),
    "gpumlib_nmf_server_create": (_i, [_c.POINTER(ServerConfig), _c.POINTER(_vp)]),
    "gpumlib_nmf_server_submit": (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _f, _d, _i, _c.POINTER(_c.c_longlong)]),
    "gpumlib_nmf_server_wait": (_i, [_vp, _c.c_longlong, _vp, _vp, _c.POINTER(JobResult)]),
    "gpumlib_nmf_server_stats_get": (_i, [_vp, _c.POINTER(ServerStats)]),
    "gpumlib_nmf_server_destroy": (None, [_vp]),
    "gpumlib_reduce_workspace_bytes": (_sz, [_l]),
    "gpumlib_reduce_sum": (_i, [_vp, _l, _f, _vp, _vp, _sz, _vp]),
    "gpumlib_reduce_min": (_i, [_vp, _l, _vp, _vp, _vp, _sz, _vp]),
    "gpumlib_reduce_max": (_i, [_vp, _l, _vp, _vp, _vp, _sz, _vp]),
    "gpumlib_nmf_rng_u64": (_u64, [_u64, _u64, _u64]),
    "gpumlib_nmf_gen_uniform_f32": (None, [_u64, _u64, _sz, _vp]),
    "gpumlib_nmf_gen_euclidean_v": (None, [_u64, _i, _i, _i, _f, _vp]),
    "gpumlib_nmf_gen_poisson_v": (None, [_u64, _i, _i, _i, _vp]),
    "gpumlib_nmf_gen_init": (None, [_u64, _i, _i, _i, _vp, _vp]),
    "gpumlib_nmf_fp8_quantize": (None, [_vp, _sz, _i, _vp]),
    "gpumlib_nmf_fp8_scale_for": (_i, [_f]),
    "gpumlib_nmf_fp8_decode": (_f, [_c.c_uint8, _i]),
}

_lib = None


class NativeLibraryMissing(RuntimeError):
    pass


def load():
    """Load the native library (once) and declare every C-ABI signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"{LIB_PATH} is missing; build it with `python gpumlib_amd/build.py` "
            "(or __graft_entry__.build()).  There is no non-native fallback.")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(err, what="gpumlib call"):
    if err != 0:
        raise RuntimeError(f"{what} failed with hipError_t {err}")
