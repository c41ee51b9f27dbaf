"""ctypes binding of libsmccat_hip.so (the C-ABI declared in include/smccat_hip.h).

The library is loaded from this directory only; a missing library is an error
(there is no CPU fallback for the engine).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsmccat_hip.so")

SMC_OK, SMC_ERR_ARG, SMC_ERR_HIP, SMC_ERR_NOMEM, SMC_ERR_STATE = 0, 1, 2, 3, 4
SMC_PREC_FP64, SMC_PREC_FP32 = 0, 1
SMC_K_EIGEN, SMC_K_PMAT, SMC_K_PRUNE, SMC_K_REDUCE, SMC_K_SMC, SMC_K_SITE = 0, 1, 2, 3, 4, 5
KERNEL_FAMILIES = ("eigen", "pmat", "prune", "reduce", "smc", "site")
SMC_PRUNE_MFMA, SMC_PRUNE_VALU, SMC_PRUNE_MFMA_LEVELS, SMC_PRUNE_MFMA_WALK = 0, 1, 2, 3
SMC_MIX_REFERENCE, SMC_MIX_LOGSPACE = 0, 1

# Every exported symbol and its (restype, argtypes).
_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int32)
_BP = ctypes.POINTER(ctypes.c_uint8)
_CP = ctypes.POINTER(ctypes.c_int8)
_FP = ctypes.POINTER(ctypes.c_float)
_LP = ctypes.POINTER(ctypes.c_int64)


class EngineConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int32), ("taxa", ctypes.c_int32), ("sites", ctypes.c_int32),
                ("num_rates", ctypes.c_int32), ("max_particles", ctypes.c_int32),
                ("max_classes", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("partial_slots", ctypes.c_int32), ("pmat_classes", ctypes.c_int64),
                ("workspaces", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class RejuvenationOptions(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("using_lg", ctypes.c_int32),
                ("invariant", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("seed", ctypes.c_uint64), ("scale_delta", ctypes.c_double),
                ("subtree_scale_delta", ctypes.c_double), ("shape_delta", ctypes.c_double),
                ("rate_delta", ctypes.c_double), ("rate_full_scale", ctypes.c_double),
                ("invar_alpha", ctypes.c_double), ("gibbs_probability", ctypes.c_double),
                ("dpp_alpha", ctypes.c_double), ("rate_cholesky", ctypes.c_void_p),
                ("allocation_proposal", ctypes.c_void_p), ("record_allocations", ctypes.c_int32),
                ("reserved2", ctypes.c_int32), ("stream_keys", ctypes.c_void_p)]


class Tuning(ctypes.Structure):
    _fields_ = [("blocks_per_cu", ctypes.c_int32), ("unroll", ctypes.c_int32 * 3),
                ("ms_per_sweep", ctypes.c_double), ("configs_tried", ctypes.c_int32),
                ("xcd_groups", ctypes.c_int32), ("unroll_runtime", ctypes.c_int32),
                ("chain_max_nodes", ctypes.c_int32)]


class SamplerOptions(ctypes.Structure):
    _fields_ = [("num_particles", ctypes.c_int32), ("rejuvenation_iterations", ctypes.c_int32),
                ("rejuvenation_threshold", ctypes.c_double), ("alg8_probability", ctypes.c_double),
                ("using_lg", ctypes.c_int32), ("invariant", ctypes.c_int32), ("cat", ctypes.c_int32),
                ("batch", ctypes.c_int32), ("max_iterations", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("seed", ctypes.c_uint64), ("dpp_alpha", ctypes.c_double),
                ("allocation_proposal", ctypes.c_void_p), ("proposal_warmup", ctypes.c_int32),
                ("train_epochs", ctypes.c_int32), ("train_lr", ctypes.c_double),
                ("on_iteration", ctypes.c_void_p), ("user", ctypes.c_void_p)]


# void (*on_iteration)(void *user, int32_t iteration, double temp, double ess)
ITERATION_CALLBACK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_double)


class SamplerResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("resamples", ctypes.c_int32),
                ("final_temp", ctypes.c_double), ("final_ess", ctypes.c_double),
                ("ess_sum", ctypes.c_double), ("seconds", ctypes.c_double),
                ("rejuvenation_seconds", ctypes.c_double), ("training_seconds", ctypes.c_double),
                ("move_stats", ctypes.c_int64 * 8), ("cat_stats", ctypes.c_int64 * 3)]


SIGNATURES = {
    "smc_engine_create": (_I, [ctypes.POINTER(EngineConfig), ctypes.POINTER(_P)]),
    "smc_engine_destroy": (_I, [_P]),
    "smc_engine_last_error": (ctypes.c_char_p, [_P]),
    "smc_set_alignment": (_I, [_P, _CP]),
    "smc_set_particles": (_I, [_P, _I, _I, _IP, _DP, _DP, _IP, _DP, _IP, _DP, _DP]),
    "smc_refresh_likelihood": (_I, [_P, _I, _I, _DP, _BP]),
    "smc_compute_next_step": (_I, [_P, _I, _DP, _DP, _DP, _DP, _I, _DP, _D]),
    "smc_compute_next_step_noalloc": (_I, [_P, _I, _DP, _DP, _DP, _DP, _I, _DP, _D]),
    "smc_systematic_resample": (_I, [_P, _I, _DP, _D, _IP]),
    "smc_discretize_gamma": (_I, [_D, _I, _DP]),
    "smc_lg_base_matrix": (None, [_DP]),
    "smc_engine_set_timing": (_I, [_P, _I]),
    "smc_engine_set_lane": (_I, [_P, _I]),
    "smc_engine_kernel_stats": (_I, [_P, _I, ctypes.POINTER(ctypes.c_int64), _DP]),
    "smc_engine_reset_stats": (_I, [_P]),
    "smc_engine_device_bytes": (ctypes.c_int64, [_P]),
    "smc_engine_set_variant": (_I, [_P, _I, _I]),
    "smc_engine_set_mixture": (_I, [_P, _I]),
    "smc_site_class_lnl": (_I, [_P, _I, _IP, _IP, _I, _IP, _DP, _D, _DP]),
    "smc_gibbs_partition_move": (_I, [_P, _I, _IP, _D, ctypes.c_uint64, _I, _D,
                                      ctypes.POINTER(ctypes.c_int64)]),
    "smc_get_particle_state": (_I, [_P, _I, _IP, _DP, _IP]),
    "smc_get_particle_params": (_I, [_P, _I, _IP, _DP, _DP, _DP, _DP]),
    "smc_mlp_num_params": (_I, []),
    "smc_mlp_create": (_I, [_P, ctypes.c_uint64, ctypes.POINTER(_P)]),
    "smc_mlp_destroy": (_I, [_P]),
    "smc_mlp_get_params": (_I, [_P, _FP]),
    "smc_mlp_set_params": (_I, [_P, _FP]),
    "smc_mlp_add_examples": (_I, [_P, _I, _FP, _IP, _IP, _LP, _FP, ctypes.c_int64]),
    "smc_mlp_collect_gibbs": (_I, [_P, _I, _IP, _D, ctypes.c_uint64, _I, _D, _LP]),
    "smc_mlp_num_examples": (ctypes.c_int64, [_P]),
    "smc_mlp_clear_examples": (_I, [_P]),
    "smc_mlp_get_example": (_I, [_P, ctypes.c_int64, _FP, _IP, _IP, _FP]),
    "smc_mlp_loss_grad": (_I, [_P, _I, _I, _DP, _FP]),
    "smc_mlp_infer": (_I, [_P, _I, _I, _I, _FP]),
    "smc_mlp_train": (_I, [_P, _I, _I, _D, ctypes.c_uint64, _DP]),
    "smc_learned_partition_move": (_I, [_P, _I, _IP, _D, ctypes.c_uint64, _I, _D, _D, _LP]),
    "smc_engine_phase_stats": (_I, [_P, _DP, _LP]),
    "smc_engine_reset_phase_stats": (_I, [_P]),
    "smc_engine_get_tuning": (_I, [_P, ctypes.POINTER(Tuning)]),
    "smc_engine_set_tuning": (_I, [_P, ctypes.POINTER(Tuning)]),
    "smc_engine_autotune": (_I, [_P, _I, _I, _I, ctypes.POINTER(Tuning)]),
    "smc_engine_set_graphs": (_I, [_P, _I]),
    "smc_engine_graph_stats": (_I, [_P, _LP, _LP]),
    "smc_ln_prior": (_I, [_P, _I, _I, _D, _DP]),
    "smc_run_sampler": (_I, [_P, ctypes.POINTER(SamplerOptions), ctypes.POINTER(SamplerResult), _DP, _DP]),
    "smc_update_params": (_I, [_P, _I, _I, _IP, _DP, _DP, _DP, _DP]),
    "smc_refresh_incremental": (_I, [_P, _I, _I, _BP, _BP, _DP, _BP]),
    "smc_accept_reject": (_I, [_P, _I, _I, _BP]),
    "smc_copy_particles": (_I, [_P, _I, _IP]),
    "smc_engine_set_patterns": (_I, [_P, _I]),
    "smc_split_posteriors": (_I, [_P, _I, _DP, ctypes.c_int64, _LP, ctypes.POINTER(ctypes.c_uint64), _DP, _DP]),
    "smc_consensus": (_I, [_P, _I, _DP, _IP, ctypes.POINTER(ctypes.c_uint64), _DP, _DP]),
    "smc_proposal_tuning": (_I, [_P, _I, _I, _DP, _DP]),
    "smc_pattern_count": (ctypes.c_int64, [_P, _I, _I]),
    "smc_rejuvenation_steps": (_I, [_P, ctypes.POINTER(RejuvenationOptions), _I, _IP, _IP]),
    "smc_rejuvenate": (_I, [_P, _I, _I, _D, ctypes.POINTER(RejuvenationOptions), _DP,
                            ctypes.POINTER(ctypes.c_int64)]),
}

_lib = None


def load():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsmccat_hip.so not built: run `python smc-cat_amd/smccat/build.py` "
                           "(or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
