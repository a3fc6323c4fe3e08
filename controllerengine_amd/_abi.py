"""ctypes mirror of include/ControllerEngine.h and the library loader.

The product path is the HIP library libControllerEngine.so built in-tree by
controllerengine_amd/build.py; loading fails loudly when it is missing.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libControllerEngine.so")

MAX_SS = 8
MAX_DS = 8
MAX_NONLIN = 4
MAX_TOTAL = 16
MAX_NL_DIM = 9
MAX_DC = 4
MAX_U = 4
MAX_NL = 8
MAX_PARAMS = 34

PRECISION_FP64 = 0
PRECISION_MIXED = 1


class CeConfig(ctypes.Structure):
    _fields_ = [
        ("statespace_dim", ctypes.c_int32), ("discrete_dim", ctypes.c_int32), ("nonlin_dim", ctypes.c_int32),
        ("n_params", ctypes.c_int32), ("input_size", ctypes.c_int32),
        ("dc_instructions", ctypes.c_int32), ("u_instructions", ctypes.c_int32),
        ("nl_instructions", ctypes.c_int32 * MAX_NONLIN),
        ("n_cores", ctypes.c_int32), ("n_rounds", ctypes.c_int32), ("serial", ctypes.c_int32),
        ("time", ctypes.c_int32), ("duration", ctypes.c_float), ("ref_step_time", ctypes.c_double),
        ("ref_amplitude", ctypes.c_float),
        ("fit_start", ctypes.c_float), ("fit_end", ctypes.c_float), ("steady_start", ctypes.c_float),
        ("bandwidth", ctypes.c_float), ("damping", ctypes.c_float), ("overshoot", ctypes.c_float),
        ("steady_state_error", ctypes.c_float), ("crossover_coeff", ctypes.c_float),
        ("target_gm", ctypes.c_float), ("target_pm", ctypes.c_float), ("target_b", ctypes.c_float),
        ("eig_fail_err", ctypes.c_float), ("unstable_err", ctypes.c_float),
        ("freq", ctypes.c_int32), ("freq_min", ctypes.c_float), ("dump_diagnostics", ctypes.c_int32),
    ]


class CePlant(ctypes.Structure):
    _fields_ = [
        ("ss_a", ctypes.c_float * (MAX_SS * MAX_SS)), ("ss_a_i", ctypes.c_float * (MAX_SS * MAX_SS)),
        ("ss_b", ctypes.c_float * MAX_SS), ("ss_c", ctypes.c_float * MAX_SS), ("ss_d", ctypes.c_float),
        ("d_a", ctypes.c_float * (MAX_DS * MAX_DS)), ("d_a_i", ctypes.c_float * (MAX_DS * MAX_DS)),
        ("d_bu", ctypes.c_float * MAX_DS), ("d_bw", ctypes.c_float * (MAX_DS * MAX_NONLIN)),
        ("d_c", ctypes.c_float * MAX_DS), ("base_params", ctypes.c_float * MAX_PARAMS),
        ("dc_info", ctypes.c_int32 * (MAX_DC * MAX_DS)), ("u_info", ctypes.c_int32 * (MAX_U * MAX_SS)),
        ("nl_info", ctypes.c_int32 * (MAX_NONLIN * MAX_NL * MAX_NL_DIM)),
    ]


class CeDerived(ctypes.Structure):
    _fields_ = [
        ("n_threads", ctypes.c_int32), ("total_dim", ctypes.c_int32), ("cuda_mx_width", ctypes.c_int32),
        ("info_mx_width", ctypes.c_int32), ("info_scoring", ctypes.c_int32), ("info_freqresp", ctypes.c_int32),
        ("info_allmargin", ctypes.c_int32), ("info_modal", ctypes.c_int32), ("fit_start_id", ctypes.c_int32),
        ("steady_start_id", ctypes.c_int32), ("fit_count", ctypes.c_int32),
        ("ts", ctypes.c_float), ("freq_max", ctypes.c_float), ("fit_t0", ctypes.c_float), ("fit_t1", ctypes.c_float),
        ("crossover", ctypes.c_float),
    ]


# exported symbols of include/ControllerEngine.h (checked by tests/test_fixtures.py)
EXPORTS = [
    "RunControllerEngine", "ce_config_parse", "ce_config_derive", "ce_plant_load_dir", "ce_engine_create",
    "ce_engine_destroy", "ce_evaluate", "ce_engine_set_diagnostics", "ce_engine_stage_times",
    "ce_engine_last_sim_jit", "ce_engine_jit_compiles", "ce_engine_set_sim_history", "ce_engine_sim_history",
    "ce_engine_last_sim_kernel", "ce_engine_last_sim_split", "ce_sim_split_slots", "ce_engine_set_profiling", "ce_engine_kernel_times", "ce_engine_sweep_counts",
    "ce_detmath_host",
    "ce_detmath_device", "ce_batch_create", "ce_batch_destroy", "ce_batch_clear", "ce_batch_add", "ce_batch_update",
    "ce_batch_offsets", "ce_batch_set_options", "ce_batch_set_plan", "ce_batch_plan", "ce_batch_run", "ce_batch_info", "ce_batch_kernel_times",
    "ce_server_create", "ce_server_destroy", "ce_server_submit", "ce_server_poll", "ce_server_drain",
    "ce_server_stats", "ce_server_reset_stats", "ce_rank", "ce_rank_noalloc", "ce_simulate_kalman",
]

_lib = None


def _declare(lib):
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    lib.RunControllerEngine.restype = ctypes.c_int
    lib.RunControllerEngine.argtypes = []
    lib.ce_config_parse.argtypes = [ctypes.c_char_p, ctypes.c_char_p, P(CeConfig), ctypes.c_char_p, ctypes.c_int]
    lib.ce_config_parse.restype = ctypes.c_int
    lib.ce_config_derive.argtypes = [P(CeConfig), P(CeDerived), ctypes.c_char_p, ctypes.c_int]
    lib.ce_config_derive.restype = ctypes.c_int
    lib.ce_plant_load_dir.argtypes = [ctypes.c_char_p, P(CeConfig), ctypes.c_int, P(CePlant), ctypes.c_char_p, ctypes.c_int]
    lib.ce_plant_load_dir.restype = ctypes.c_int
    lib.ce_engine_create.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.ce_engine_create.restype = vp
    lib.ce_engine_destroy.argtypes = [vp]
    lib.ce_engine_destroy.restype = None
    lib.ce_evaluate.argtypes = [vp, P(CeConfig), P(CePlant), ctypes.c_int, vp, vp, vp, vp, ctypes.c_int, vp,
                                ctypes.c_char_p, ctypes.c_int]
    lib.ce_evaluate.restype = ctypes.c_int
    lib.ce_engine_set_diagnostics.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.ce_engine_set_diagnostics.restype = ctypes.c_int
    lib.ce_engine_stage_times.argtypes = [vp, P(ctypes.c_double), ctypes.c_int]
    lib.ce_engine_stage_times.restype = ctypes.c_int
    lib.ce_engine_last_sim_jit.argtypes = [vp]
    lib.ce_engine_last_sim_jit.restype = ctypes.c_int
    lib.ce_engine_jit_compiles.argtypes = [vp]
    lib.ce_engine_jit_compiles.restype = ctypes.c_longlong
    lib.ce_engine_last_sim_kernel.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    lib.ce_engine_last_sim_kernel.restype = ctypes.c_int
    lib.ce_engine_last_sim_split.argtypes = [vp]
    lib.ce_engine_last_sim_split.restype = ctypes.c_int
    lib.ce_sim_split_slots.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.ce_sim_split_slots.restype = ctypes.c_int
    lib.ce_engine_set_sim_history.argtypes = [vp, ctypes.c_int]
    lib.ce_engine_set_sim_history.restype = ctypes.c_int
    lib.ce_engine_sim_history.argtypes = [vp, P(ctypes.c_double), ctypes.c_int]
    lib.ce_engine_sim_history.restype = ctypes.c_int
    lib.ce_engine_set_profiling.argtypes = [vp, ctypes.c_int]
    lib.ce_engine_set_profiling.restype = ctypes.c_int
    lib.ce_engine_kernel_times.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, P(ctypes.c_double), ctypes.c_int]
    lib.ce_engine_kernel_times.restype = ctypes.c_int
    lib.ce_engine_sweep_counts.argtypes = [vp, P(ctypes.c_int), P(ctypes.c_int)]
    lib.ce_engine_sweep_counts.restype = ctypes.c_int
    lib.ce_detmath_host.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int]
    lib.ce_detmath_host.restype = ctypes.c_int
    lib.ce_detmath_device.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int, vp]
    lib.ce_detmath_device.restype = ctypes.c_int
    lib.ce_rank.argtypes = [P(CeConfig), vp, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, ctypes.c_char_p, ctypes.c_int]
    lib.ce_rank.restype = ctypes.c_int
    lib.ce_simulate_kalman.argtypes = [vp, P(CeConfig), P(CePlant), ctypes.c_int, vp, ctypes.c_int, vp, vp, vp,
                                       ctypes.c_char_p, ctypes.c_int]
    lib.ce_simulate_kalman.restype = ctypes.c_int
    lib.ce_rank_noalloc.argtypes = [P(CeConfig), vp, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, ctypes.c_char_p,
                                    ctypes.c_int]
    lib.ce_rank_noalloc.restype = ctypes.c_int
    lib.ce_batch_create.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.ce_batch_create.restype = vp
    lib.ce_batch_destroy.argtypes = [vp]
    lib.ce_batch_destroy.restype = None
    lib.ce_batch_set_plan.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.ce_batch_set_plan.restype = ctypes.c_int
    lib.ce_batch_plan.argtypes = [vp, P(ctypes.c_int)]
    lib.ce_batch_plan.restype = ctypes.c_int
    lib.ce_batch_clear.argtypes = [vp]
    lib.ce_batch_clear.restype = ctypes.c_int
    lib.ce_batch_add.argtypes = [vp, P(CeConfig), P(CePlant), ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.ce_batch_add.restype = ctypes.c_int
    lib.ce_batch_update.argtypes = [vp, ctypes.c_int, P(CeConfig), P(CePlant), ctypes.c_char_p, ctypes.c_int]
    lib.ce_batch_update.restype = ctypes.c_int
    lib.ce_batch_offsets.argtypes = [vp, ctypes.c_int, P(ctypes.c_longlong), P(ctypes.c_longlong)]
    lib.ce_batch_offsets.restype = ctypes.c_int
    lib.ce_batch_set_options.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.ce_batch_set_options.restype = ctypes.c_int
    lib.ce_batch_run.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp, vp, ctypes.c_char_p, ctypes.c_int]
    lib.ce_batch_run.restype = ctypes.c_int
    lib.ce_batch_info.argtypes = [vp, P(ctypes.c_int), P(ctypes.c_int), P(ctypes.c_int)]
    lib.ce_batch_info.restype = ctypes.c_int
    lib.ce_batch_kernel_times.argtypes = [vp, ctypes.c_char_p, ctypes.c_int, P(ctypes.c_double), ctypes.c_int]
    lib.ce_batch_kernel_times.restype = ctypes.c_int
    lib.ce_server_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_char_p, ctypes.c_int]
    lib.ce_server_create.restype = vp
    lib.ce_server_destroy.argtypes = [vp]
    lib.ce_server_destroy.restype = None
    lib.ce_server_submit.argtypes = [vp, P(CeConfig), P(CePlant), vp, vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    lib.ce_server_submit.restype = ctypes.c_longlong
    lib.ce_server_poll.argtypes = [vp, ctypes.c_longlong, vp, vp, ctypes.c_char_p, ctypes.c_int]
    lib.ce_server_poll.restype = ctypes.c_int
    lib.ce_server_drain.argtypes = [vp, ctypes.c_double]
    lib.ce_server_drain.restype = ctypes.c_int
    lib.ce_server_stats.argtypes = [vp, P(ctypes.c_double), ctypes.c_int]
    lib.ce_server_stats.restype = ctypes.c_int
    lib.ce_server_reset_stats.argtypes = [vp]
    lib.ce_server_reset_stats.restype = ctypes.c_int
    return lib


def lib():
    """The HIP engine library; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libControllerEngine.so is not built: run python controllerengine_amd/build.py "
                               "(or __graft_entry__.build())")
        # torch carries its own copy of the HIP runtime.  Loaded first, it provides the
        # libamdhip64 this library links against; loaded after this library, it would bring a
        # second runtime into the process, and whichever initialises second finds no device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = _declare(ctypes.CDLL(LIB_PATH))
    return _lib


def errbuf():
    return ctypes.create_string_buffer(1024)
