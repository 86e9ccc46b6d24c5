"""ctypes binding of ``libfagra_gpu.so`` (declarations in include/fagra_gpu.h)."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfagra_gpu.so")

OK, INVALID_OPTIONS, LINEAR_SOLVE_FAILED, NO_CONVERGENCE, NO_PROGRESS = 0, 1, 2, 3, 4
INVALID_EVALUATION, DIMENSION_MISMATCH, INVALID_EMISSION, UNSUPPORTED_STRUCTURE = 5, 6, 7, 8
INVALID_INPUT, DEVICE_ERROR = 9, 100
DUPLICATE_COVARIANCE_BLOCK, SINGULAR_INFORMATION, INVALID_INFORMATION, INVALID_RANK_TOLERANCE = 10, 11, 12, 13
UNKNOWN_KEY = 14
FP64, FP32 = 0, 1
PRECOND_BLOCK_JACOBI, PRECOND_LEARNED = 0, 1
CG_NONE, CG_CONVERGED, CG_ITERATION_LIMIT, CG_NOT_POSITIVE_DEFINITE, CG_NON_FINITE = -1, 0, 1, 2, 3
TERM_NONE, TERM_NO_VARIABLES, TERM_GRADIENT, TERM_STEP, TERM_COST = -1, 0, 1, 2, 3
REJ_NONE, REJ_LINEAR_SOLVE, REJ_PREDICTION, REJ_INVALID_EVALUATION, REJ_POOR_AGREEMENT = -1, 0, 1, 2, 3

STATUS_NAMES = {OK: "Ok", INVALID_OPTIONS: "InvalidOptions", LINEAR_SOLVE_FAILED: "LinearSolveFailed",
                NO_CONVERGENCE: "NoConvergence", NO_PROGRESS: "NoProgress",
                INVALID_EVALUATION: "InvalidEvaluation", DIMENSION_MISMATCH: "DimensionMismatch",
                INVALID_EMISSION: "InvalidEmission", UNSUPPORTED_STRUCTURE: "UnsupportedStructure",
                INVALID_INPUT: "InvalidInput", DUPLICATE_COVARIANCE_BLOCK: "DuplicateCovarianceBlock",
                SINGULAR_INFORMATION: "SingularInformation", INVALID_INFORMATION: "InvalidInformation",
                INVALID_RANK_TOLERANCE: "InvalidRankTolerance", UNKNOWN_KEY: "Unknown", DEVICE_ERROR: "DeviceError"}

_sz = ctypes.c_size_t
_d = ctypes.c_double
_dp = ctypes.POINTER(ctypes.c_double)
_i64p = ctypes.POINTER(ctypes.c_int64)
_szp = ctypes.POINTER(ctypes.c_size_t)
_fp = ctypes.POINTER(ctypes.c_float)
_VP = ctypes.c_void_p


class SchurDiagnostics(ctypes.Structure):
    _fields_ = [("termination", ctypes.c_int), ("iterations", _sz), ("damping", _d),
                ("estimated_relative_residual", _d), ("reduced_relative_residual", _d),
                ("scaled_relative_residual", _d), ("original_relative_residual", _d),
                ("used_fallback", ctypes.c_int)]


class SchurTimings(ctypes.Structure):
    _fields_ = [("preparation", _d), ("factorization", _d), ("linear_solve", _d), ("products", _d),
                ("preconditioner", _d), ("verification", _d), ("linearization", _d), ("cost", _d),
                ("matvecs", _sz), ("matvec_kernel_seconds", _d)]


ON_SOLVE = ctypes.CFUNCTYPE(None, ctypes.POINTER(SchurDiagnostics), _VP)


class SchurConfig(ctypes.Structure):
    _fields_ = [("max_iterations", _sz), ("relative_tolerance", _d), ("collect_timings", ctypes.c_int),
                ("precision", ctypes.c_int), ("preconditioner", ctypes.c_int), ("fallback_factor", _d),
                ("fallback_iterations", _sz), ("on_solve", ON_SOLVE), ("on_solve_user", _VP)]


LSMR_NONE, LSMR_CONVERGED, LSMR_ITERATION_LIMIT, LSMR_NON_FINITE = -1, 0, 1, 2


class LsmrDiagnostics(ctypes.Structure):
    _fields_ = [("termination", ctypes.c_int), ("iterations", _sz), ("finite_step", ctypes.c_int),
                ("estimated_absolute_residual", _d), ("estimated_relative_residual", _d),
                ("original_absolute_residual", _d), ("original_relative_residual", _d),
                ("scaled_absolute_residual", _d), ("scaled_relative_residual", _d),
                ("min_scale", _d), ("max_scale", _d), ("damping", _d), ("block_fallbacks", _sz)]


class LsmrTimings(ctypes.Structure):
    _fields_ = [("preparation", _d), ("factorization", _d), ("linear_solve", _d), ("forward", _d),
                ("transpose", _d), ("preconditioner", _d), ("residual_checks", _d)]


ON_LSMR_SOLVE = ctypes.CFUNCTYPE(None, ctypes.POINTER(LsmrDiagnostics), _VP)


class LsmrConfig(ctypes.Structure):
    _fields_ = [("max_iterations", _sz), ("relative_tolerance", _d), ("diagonal_preconditioning", ctypes.c_int),
                ("block_preconditioning", ctypes.c_int), ("collect_timings", ctypes.c_int),
                ("precision", ctypes.c_int), ("on_solve", ON_LSMR_SOLVE), ("on_solve_user", _VP)]


class CovarianceOptions(ctypes.Structure):
    _fields_ = [("relative_pivot_tolerance", _d)]


class BalSizes(ctypes.Structure):
    _fields_ = [("n_cameras", _sz), ("n_points", _sz), ("n_observations", _sz)]


class OptimizeOptions(ctypes.Structure):
    _fields_ = [("max_iterations", _sz), ("gradient_tolerance", _d), ("step_tolerance", _d),
                ("cost_tolerance", _d)]


class LmOptions(ctypes.Structure):
    _fields_ = [("initial_damping", _d), ("min_damping", _d), ("max_damping", _d), ("min_column_norm", _d),
                ("acceptance_threshold", _d), ("max_trials", _sz)]


class LmOutcome(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("termination", ctypes.c_int), ("iterations", _sz),
                ("initial_cost", _d), ("final_cost", _d), ("linearizations", _sz), ("attempts", _sz),
                ("rejected_steps", _sz), ("accepted_steps", _sz), ("cost_evaluations", _sz),
                ("damping", _d), ("cost", _d), ("gradient_norm", _d), ("step_norm", _d),
                ("last_rejection", ctypes.c_int), ("linear_solves", _sz), ("linear_iterations", _sz)]


ON_ACCEPT = ctypes.CFUNCTYPE(None, ctypes.POINTER(LmOutcome), _VP)
FN_PREPARE = ctypes.CFUNCTYPE(ctypes.c_int, _VP)
FN_COST = ctypes.CFUNCTYPE(ctypes.c_int, _VP, ctypes.c_int, _dp)
FN_LIN = ctypes.CFUNCTYPE(ctypes.c_int, _VP)
FN_GN = ctypes.CFUNCTYPE(ctypes.c_int, _VP, _dp)
FN_PD = ctypes.CFUNCTYPE(ctypes.c_int, _VP, _d)
FN_SD = ctypes.CFUNCTYPE(ctypes.c_int, _VP, _d, _dp, _dp)
FN_S = ctypes.CFUNCTYPE(ctypes.c_int, _VP, _dp)
FN_STAGE = ctypes.CFUNCTYPE(ctypes.c_int, _VP)
FN_VOID = ctypes.CFUNCTYPE(None, _VP)
FN_WORK = ctypes.CFUNCTYPE(None, _VP, _szp, _szp)


class Problem(ctypes.Structure):
    _fields_ = [("ctx", _VP), ("dimension", _sz), ("prepare", FN_PREPARE), ("cost", FN_COST),
                ("linearize", FN_LIN), ("gradient_norm", FN_GN), ("prepare_damping", FN_PD),
                ("solve_damped", FN_SD), ("solve", FN_S), ("stage", FN_STAGE), ("accept", FN_VOID),
                ("reject", FN_VOID), ("linear_work", FN_WORK)]


# Every exported symbol with its signature; tests check the library exports exactly these.
SIGNATURES = {
    "fg_ba_create": (ctypes.c_int, [_sz, _sz, _sz, _dp, _dp, _i64p, _i64p, _dp, _d, ctypes.c_int,
                                    ctypes.POINTER(_VP)]),
    "fg_ba_destroy": (None, [_VP]),
    "fg_ba_layout": (ctypes.c_int, [_VP, _i64p, _i64p, _i64p, _i64p]),
    "fg_ba_get_state": (ctypes.c_int, [_VP, _dp, _dp]),
    "fg_ba_set_state": (ctypes.c_int, [_VP, _dp, _dp]),
    "fg_ba_cost": (ctypes.c_int, [_VP, ctypes.c_int, _dp]),
    "fg_ba_stage": (ctypes.c_int, [_VP, _dp]),
    "fg_ba_accept": (None, [_VP]),
    "fg_ba_reject": (None, [_VP]),
    "fg_schur_config_default": (ctypes.c_int, [ctypes.POINTER(SchurConfig), ctypes.c_int]),
    "fg_schur_create": (ctypes.c_int, [ctypes.POINTER(SchurConfig), ctypes.POINTER(_VP)]),
    "fg_schur_destroy": (None, [_VP]),
    "fg_schur_set_config": (ctypes.c_int, [_VP, ctypes.POINTER(SchurConfig)]),
    "fg_schur_prepare": (ctypes.c_int, [_VP, _VP, _sz]),
    "fg_schur_clear": (None, [_VP]),
    "fg_schur_linearize": (ctypes.c_int, [_VP]),
    "fg_schur_gradient_norm": (ctypes.c_int, [_VP, _dp]),
    "fg_schur_prepare_damping": (ctypes.c_int, [_VP, _d]),
    "fg_schur_solve_damped": (ctypes.c_int, [_VP, _d, ctypes.POINTER(_dp), _dp]),
    "fg_schur_solve": (ctypes.c_int, [_VP, ctypes.POINTER(_dp)]),
    "fg_schur_diagnostics_get": (None, [_VP, ctypes.POINTER(SchurDiagnostics)]),
    "fg_schur_statistics": (None, [_VP, _szp, _szp]),
    "fg_schur_timings_get": (None, [_VP, ctypes.POINTER(SchurTimings)]),
    "fg_schur_model": (ctypes.c_int, [_VP, _dp, _dp, _dp]),
    "fg_schur_camera_gram": (ctypes.c_int, [_VP, _dp]),
    "fg_schur_reduced_apply": (ctypes.c_int, [_VP, _d, _dp, _dp]),
    "fg_schur_time_products": (ctypes.c_int, [_VP, _d, _dp, ctypes.c_int, _dp, _dp]),
    "fg_schur_record_mismatches": (ctypes.c_int, [_VP]),
    "fg_schur_jacobian_apply": (ctypes.c_int, [_VP, _dp, _dp]),
    "fg_schur_jacobian_transpose_apply": (ctypes.c_int, [_VP, _dp, _dp]),
    "fg_schur_damped_apply": (ctypes.c_int, [_VP, _d, ctypes.c_int, _dp, _dp]),
    "fg_schur_damped_transpose_apply": (ctypes.c_int, [_VP, _d, ctypes.c_int, _dp, _dp]),
    "fg_optimize_options_default": (ctypes.c_int, [ctypes.POINTER(OptimizeOptions)]),
    "fg_lm_options_default": (ctypes.c_int, [ctypes.POINTER(LmOptions)]),
    "fg_lsmr_config_default": (ctypes.c_int, [ctypes.POINTER(LsmrConfig), ctypes.c_int]),
    "fg_lsmr_create": (ctypes.c_int, [ctypes.POINTER(LsmrConfig), ctypes.POINTER(_VP)]),
    "fg_lsmr_destroy": (None, [_VP]),
    "fg_lsmr_set_config": (ctypes.c_int, [_VP, ctypes.POINTER(LsmrConfig)]),
    "fg_lsmr_prepare": (ctypes.c_int, [_VP, _VP, _sz]),
    "fg_lsmr_clear": (None, [_VP]),
    "fg_lsmr_linearize": (ctypes.c_int, [_VP]),
    "fg_lsmr_gradient_norm": (ctypes.c_int, [_VP, _dp]),
    "fg_lsmr_prepare_damping": (ctypes.c_int, [_VP, _d]),
    "fg_lsmr_solve_damped": (ctypes.c_int, [_VP, _d, ctypes.POINTER(_dp), _dp]),
    "fg_lsmr_solve": (ctypes.c_int, [_VP, ctypes.POINTER(_dp)]),
    "fg_lsmr_diagnostics_get": (None, [_VP, ctypes.POINTER(LsmrDiagnostics)]),
    "fg_lsmr_statistics": (None, [_VP, _szp, _szp]),
    "fg_lsmr_timings_get": (None, [_VP, ctypes.POINTER(LsmrTimings)]),
    "fg_lm_solve_batch_lsmr": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(OptimizeOptions), ctypes.POINTER(LmOptions),
                                              ON_ACCEPT, _VP, ctypes.POINTER(LmOutcome)]),
    "fg_gn_solve_batch_lsmr": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(OptimizeOptions),
                                              ctypes.POINTER(LmOutcome)]),
    "fg_gschur_create": (ctypes.c_int, [ctypes.POINTER(SchurConfig), _sz, _sz, _sz, ctypes.POINTER(_VP)]),
    "fg_gschur_destroy": (None, [_VP]),
    "fg_gschur_set_config": (ctypes.c_int, [_VP, ctypes.POINTER(SchurConfig)]),
    "fg_gschur_prepare": (ctypes.c_int, [_VP, _sz]),
    "fg_gschur_clear": (None, [_VP]),
    "fg_gschur_accumulate": (ctypes.c_int, [_VP, _sz, _sz, _dp, _sz, _szp, _szp, _dp]),
    "fg_gschur_gradient_norm": (ctypes.c_int, [_VP, _dp]),
    "fg_gschur_prepare_damping": (ctypes.c_int, [_VP, _d]),
    "fg_gschur_solve_damped": (ctypes.c_int, [_VP, _d, ctypes.POINTER(_dp), _dp]),
    "fg_gschur_solve": (ctypes.c_int, [_VP, ctypes.POINTER(_dp)]),
    "fg_gschur_diagnostics_get": (None, [_VP, ctypes.POINTER(SchurDiagnostics)]),
    "fg_gschur_statistics": (None, [_VP, _szp, _szp]),
    "fg_gschur_last_step_norm": (_d, [_VP]),
    "fg_covariance_options_default": (ctypes.c_int, [ctypes.POINTER(CovarianceOptions)]),
    "fg_dense_covariance": (ctypes.c_int, [_dp, _sz, _szp, _sz, ctypes.POINTER(CovarianceOptions), _dp]),
    "fg_dense_covariance_noalloc": (ctypes.c_int, [_dp, _sz, _szp, _sz, ctypes.POINTER(CovarianceOptions), _dp]),
    "fg_schur_covariance": (ctypes.c_int, [_VP, _szp, _szp, _sz, ctypes.POINTER(CovarianceOptions), _dp]),
    "fg_gschur_covariance": (ctypes.c_int, [_VP, _szp, _szp, _sz, ctypes.POINTER(CovarianceOptions), _dp]),
    "fg_bal_scan": (ctypes.c_int, [ctypes.c_char_p, _sz, ctypes.POINTER(BalSizes), ctypes.c_char_p, _sz]),
    "fg_bal_parse": (ctypes.c_int, [ctypes.c_char_p, _sz, ctypes.POINTER(BalSizes), _dp, _dp, _i64p, _i64p, _dp,
                                    ctypes.c_char_p, _sz]),
    "fg_bal_format": (ctypes.c_int, [ctypes.POINTER(BalSizes), _dp, _dp, _i64p, _i64p, _dp, ctypes.c_char_p, _sz,
                                     _szp]),
    "fg_lm_solve_batch": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(OptimizeOptions), ctypes.POINTER(LmOptions),
                                         ON_ACCEPT, _VP, ctypes.POINTER(LmOutcome)]),
    "fg_gn_solve_batch": (ctypes.c_int, [_VP, _VP, ctypes.POINTER(OptimizeOptions), ctypes.POINTER(LmOutcome)]),
    "fg_lm_optimize": (None, [ctypes.POINTER(Problem), ctypes.POINTER(OptimizeOptions), ctypes.POINTER(LmOptions),
                              ON_ACCEPT, _VP, ctypes.POINTER(LmOutcome)]),
    "fg_gn_optimize": (None, [ctypes.POINTER(Problem), ctypes.POINTER(OptimizeOptions),
                              ctypes.POINTER(LmOutcome)]),
    "fg_device_info": (ctypes.c_int, [ctypes.c_char_p, _sz, ctypes.POINTER(ctypes.c_int), _szp]),
    "fg_mlp_param_count": (_sz, []),
    "fg_mlp_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.POINTER(_VP)]),
    "fg_mlp_destroy": (None, [_VP]),
    "fg_mlp_set_params": (ctypes.c_int, [_VP, _fp]),
    "fg_mlp_set_input_range": (ctypes.c_int, [_VP, _fp, _fp]),
    "fg_mlp_get_params": (ctypes.c_int, [_VP, _fp]),
    "fg_mlp_forward": (ctypes.c_int, [_VP, _sz, _fp, _fp]),
    "fg_mlp_loss_grad": (ctypes.c_int, [_VP, _sz, _fp, _fp, _dp, _fp]),
    "fg_mlp_train": (ctypes.c_int, [_VP, _sz, _fp, _fp, _sz, _sz, _d, ctypes.c_uint64, _dp]),
    "fg_schur_set_mlp": (ctypes.c_int, [_VP, _VP]),
    "fg_schur_capture": (ctypes.c_int, [_VP, _d, _fp, _fp, ctypes.POINTER(ctypes.c_int)]),
    "fg_status_string": (ctypes.c_char_p, [ctypes.c_int]),
}

_lib = None


def load():
    """Load the HIP library; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m fagra_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class FagraError(RuntimeError):
    def __init__(self, status, where=""):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}")


def check(status, where=""):
    if status != OK:
        raise FagraError(status, where)
    return status
