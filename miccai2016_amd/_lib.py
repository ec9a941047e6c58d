"""ctypes binding of libmiccai_hip.so (declared in include/miccai_hip.h).

The library is the only compute path: if it is missing or fails to load, the
import of any processing module raises - there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmiccai_hip.so")

MH_OK = 0
MH_ERR_ARG = -1
MH_ERR_HIP = -2
MH_STOPPED_NAN = 1
MH_DEFER = 2

_lib = None

vp = C.c_void_p
i64 = C.c_int64
i32 = C.c_int
f32 = C.c_float
I64x3 = i64 * 3
F32x3 = C.c_float * 3
I32x3 = C.c_int * 3


class GvfParams(C.Structure):
    _fields_ = [("mu", f32), ("dt", f32), ("rss", f32 * 3)]


class CmfParams(C.Structure):
    _fields_ = [("rcc", f32), ("rs", f32 * 3), ("ss", f32 * 3), ("sqev", f32 * 3),
                ("const_evals", i32), ("num_iter", i32), ("update", i32), ("error_bound", f32)]


class CmfResult(C.Structure):
    _fields_ = [("iterations", i32), ("stopped", i32), ("ncrit", i32), ("last_crit", f32)]


class CmfResident(C.Structure):
    _fields_ = [("dims", i64 * 3), ("u", vp), ("eigs", vp), ("alpha", vp), ("cs", vp), ("ct", vp),
                ("work", vp), ("params", CmfParams), ("iter", i32), ("cur", i32), ("swept", i32),
                ("method", i32), ("zchunk", i32)]


class CmfJob(C.Structure):
    _fields_ = [("dims", i64 * 3), ("u", vp), ("eigs", vp), ("alpha", vp), ("cs", vp), ("ct", vp),
                ("state", vp), ("params", CmfParams), ("iter", i32), ("stopped", i32)]


# name -> (restype, argtypes); every symbol declared in include/miccai_hip.h
SIGNATURES = {
    "mh_create": (i32, [i32, C.POINTER(vp)]),
    "mh_destroy": (i32, [vp]),
    "mh_last_error": (C.c_char_p, [vp]),
    "mh_version": (i32, []),
    "mh_fill": (i32, [vp, vp, i64, f32, vp]),
    "mh_reduce_sum_abs": (i32, [vp, vp, i64, vp, vp]),
    "mh_reduce_max": (i32, [vp, vp, i64, vp, vp]),
    "mh_gauss_filter": (i32, [vp, vp, vp, vp, I64x3, vp, i32, vp, i32, vp, i32, vp]),
    "mh_hessian_work_floats": (i64, [I64x3]),
    "mh_gauss_hessian": (i32, [vp, vp, I64x3, vp, I32x3, vp, vp, vp]),
    "mh_eigs": (i32, [vp, vp, vp, i64, i32, vp, vp]),
    "mh_eig_norm_max": (i32, [vp, vp, i64, vp, vp]),
    "mh_update_vesselness": (i32, [vp, vp, vp, vp, i64, f32, f32, f32, vp, i32, vp]),
    "mh_detect_ridges": (i32, [vp, vp, vp, vp, I64x3, f32, vp]),
    "mh_main_dirs": (i32, [vp, vp, vp, i64, vp]),
    "mh_nonridge_work_bytes": (i64, [I64x3]),
    "mh_nonridge_signs": (i32, [vp, vp, I64x3, vp, vp, i64, F32x3, vp, i64, vp]),
    "mh_complete_frames": (i32, [vp, vp, vp, i64, vp]),
    "mh_hysteresis_host": (i32, [vp, vp, vp, I64x3]),
    "mh_ridge_segments_work_bytes": (i64, [I64x3]),
    "mh_ridge_segments": (i32, [vp, vp, vp, f32, i32, I64x3, vp, vp, vp, i64, vp]),
    "mh_spanning_tree_host": (i32, [vp, i64, C.c_double, vp]),
    "mh_ridge_tree_host": (i32, [vp, i64, I64x3, vp, vp]),
    "mh_adjust_ridge_device": (i32, [vp, vp, vp, vp, I64x3, vp, i32, vp, vp, vp]),
    "mh_ridge_forest_host": (i64, [vp, i64, I64x3, vp, vp]),
    "mh_ridge_tree_from_forest_host": (i32, [vp, i64, I64x3, vp, vp, i64, vp]),
    "mh_forest_pieces_host": (i32, [vp, i64, vp, i64, vp, vp, vp]),
    "mh_gvf_set_sweep": (i32, [vp, i32]),
    "mh_tree_signs_host": (i32, [vp, i64, vp, i64, i64, vp]),
    "mh_gvf_response": (i32, [vp, vp, vp, i64, vp]),
    "mh_gvf_work_floats": (i64, [I64x3]),
    "mh_gvf": (i32, [vp, vp, vp, vp, vp, I64x3, C.POINTER(GvfParams), i32, vp, vp, vp]),
    "mh_gvf_normalize": (i32, [vp, vp, i64, vp]),
    "mh_cmf_workspace_bytes": (i64, [i64]),
    "mh_cmf_solve": (i32, [vp, vp, vp, vp, vp, vp, vp, I64x3, C.POINTER(CmfParams), vp, i64,
                           C.POINTER(CmfResult), vp, vp]),
    "mh_cmf_solve_noalloc": (i32, [vp, vp, vp, vp, vp, vp, vp, I64x3, C.POINTER(CmfParams),
                                   C.POINTER(CmfResult), vp, vp]),
    "mh_cmf_batched_workspace_bytes": (i64, [i64, i32]),
    "mh_cmf_solve_batched": (i32, [vp, vp, i32, vp, vp, vp, vp, vp, vp, I64x3,
                                   C.POINTER(CmfParams), vp, i64, C.POINTER(CmfResult), vp]),
    "mh_terminal_capacities": (i32, [vp, vp, i64, f32, f32, vp, vp, vp]),
    "mh_grad_mag_sq": (i32, [vp, vp, I64x3, vp, vp]),
    "mh_nonterminal_weight": (i32, [vp, vp, i64, f32, f32, vp, vp]),
    "mh_normalize_work_floats": (i64, [i64]),
    "mh_minmax_normalize": (i32, [vp, vp, i64, vp, vp, vp]),
    "mh_weighted_dirs": (i32, [vp, vp, vp, i64, vp, vp]),
    "mh_soa_to_aos3": (i32, [vp, vp, i64, vp, vp]),
    "mh_cmf_batched_set_schedule": (i32, [vp, i32, i32]),
    "mh_cmf_set_sweep": (i32, [vp, i32, i32, i32]),
    "mh_cmf_batched_set_sweep": (i32, [vp, i32]),
    "mh_cmf_sweep_choice": (i32, [vp, I64x3, vp, vp]),
    "mh_cmf_batched_sweep_choice": (i32, [vp, I64x3, vp, vp]),
    "mh_cmf_set_graphs": (i32, [vp, i32]),
    "mh_initial_u_work_floats": (i64, [i64]),
    "mh_initial_u": (i32, [vp, vp, i64, vp, vp, vp]),
    "mh_dice_work_bytes": (i64, [i64, i32]),
    "mh_largest_component_work_bytes": (i64, [I64x3]),
    "mh_cmf_resident_init": (i32, [vp, vp, vp]),
    "mh_cmf_resident_step": (i32, [vp, vp, i32, i32, vp]),
    "mh_cmf_resident_status": (i32, [vp, vp, vp, vp]),
    "mh_cmf_resident_finish": (i32, [vp, vp, vp, vp, vp]),
    "mh_largest_component_dice": (i32, [vp, vp, f32, vp, I64x3, vp, vp, i64, vp, vp]),
    "mh_dice_counts": (i32, [vp, vp, i32, i64, vp, vp, i64, vp, vp]),
    "mh_helix_phantom": (i32, [vp, I64x3, f32, f32, f32, f32, C.c_uint64, vp, vp, vp, vp]),
    "mh_cmf_job_state_bytes": (i64, [i64, C.POINTER(CmfParams)]),
    "mh_cmf_jobs_init": (i32, [vp, C.POINTER(CmfJob), i32, vp]),
    "mh_cmf_jobs_step": (i32, [vp, C.POINTER(CmfJob), i32, i32, vp]),
    "mh_cmf_ragged_workspace_bytes": (i64, [i64, i32]),
    "mh_cmf_solve_ragged": (i32, [vp, vp, vp, vp, vp, vp, i32, vp, vp, C.POINTER(CmfParams), vp,
                                  i64, C.POINTER(CmfResult), vp]),
}


def lib():
    """Load libmiccai_hip.so once; raise if it is absent or incomplete."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libmiccai_hip.so is not built; run `python -m miccai2016_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)  # AttributeError names a missing export
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def dims3(shape):
    return I64x3(*[int(s) for s in shape])
