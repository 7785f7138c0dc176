"""ctypes mirror of include/rewl_c.h and the loader of the HIP engine.

The engine library is built in-tree (``python -m rewl_simulator_amd.build``);
loading fails loudly when it is missing -- there is no CPU fallback.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
ENGINE_SO = os.path.join(PKG, "librewl_hip.so")

REWL_OK = 0
REWL_ERR_INVALID = -1
REWL_ERR_DEVICE = -2
REWL_ERR_NONFINITE = -3
REWL_ERR_ITER_CAP = -4
REWL_ERR_NO_DEVICE = -5
REWL_ERR_REPLAY = -6
REWL_ERR_RANGE = -7

ISING2D = 0
ASHKIN_TELLER2D = 1

FLAT_ALL_VISITED = 0
FLAT_RATIO = 1
FLAT_MIN_OVER_MEAN = 2
FLAT_MIN_OVER_MEAN_VISITED = 3

SCHED_HALVE = 0
SCHED_ONE_OVER_T = 1

RNG_PHILOX = 0
RNG_REPLAY = 1


class SystemDesc(C.Structure):
    _fields_ = [
        ("model", C.c_int32),
        ("L", C.c_int32),
        ("J", C.c_double),
        ("K", C.c_double),
        ("h", C.c_double),
        ("field", C.POINTER(C.c_double)),
        ("e_min", C.c_double),
        ("e_max", C.c_double),
        ("bin_size", C.c_double),
        ("num_windows", C.c_int32),
        ("replicas_per_window", C.c_int32),
        ("overlap", C.c_double),
        ("seed", C.c_uint64),
        ("exclude_bins", C.POINTER(C.c_uint8)),
        ("walker_seeds", C.POINTER(C.c_uint64)),
    ]


class Params(C.Structure):
    _fields_ = [
        ("flatness_mode", C.c_int32),
        ("flatness", C.c_double),
        ("schedule", C.c_int32),
        ("lnf_initial", C.c_double),
        ("lnf_final", C.c_double),
        ("sweeps_per_check", C.c_int32),
        ("sweeps_per_exchange", C.c_int32),
        ("reduce_logdos", C.c_int32),
        ("rng_mode", C.c_int32),
        ("walkers_per_group", C.c_int32),
        ("exchange", C.c_int32),
        ("producers", C.c_int32),
        ("adjust_mode", C.c_int32),
        ("average_mode", C.c_int32),
        ("sweep_kernel", C.c_int32),
        ("observables", C.c_int32),
        ("densities", C.c_int32),
        ("retire", C.c_int32),
    ]


class AnnealParams(C.Structure):
    _fields_ = [
        ("initial_temperature", C.c_double),
        ("final_temperature", C.c_double),
        ("num_temperatures", C.c_int32),
        ("block_size", C.c_int32),
        ("stdev_tolerance", C.c_double),
        ("frozen_criterion", C.c_double),
        ("max_blocks", C.c_int32),
        ("chains", C.c_int32),
    ]


class WindowInfo(C.Structure):
    _fields_ = [
        ("system", C.c_int32),
        ("window", C.c_int32),
        ("first_walker", C.c_int32),
        ("num_bins", C.c_int32),
        ("e_min", C.c_double),
        ("e_max", C.c_double),
        ("bin_size", C.c_double),
        ("lnf", C.c_double),
        ("iteration", C.c_int64),
        ("sweeps_in_iteration", C.c_int64),
        ("done", C.c_int32),
        ("one_over_t", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("sweeps", C.c_uint64),
        ("attempts", C.c_uint64),
        ("exchange_calls", C.c_uint64),
        ("exchanges_accepted", C.c_uint64),
        ("checks", C.c_uint64),
        ("flat_events", C.c_uint64),
        ("all_done", C.c_int32),
        ("num_walkers", C.c_int32),
        ("kernel_ms_sweep", C.c_double),
        ("kernel_ms_other", C.c_double),
        ("sweep_launches", C.c_uint64),
    ]


# Every function the header declares, with its argtypes.
_P = C.c_void_p
class MlpTrainParams(C.Structure):
    _fields_ = [
        ("steps", C.c_int64),
        ("batch", C.c_int64),
        ("lr", C.c_double),
        ("lr_final", C.c_double),
        ("beta1", C.c_double),
        ("beta2", C.c_double),
        ("eps", C.c_double),
        ("seed", C.c_uint64),
        ("warmup", C.c_int64),
        ("reserved", C.c_int64 * 3),
    ]


MLP_NUM_PARAMS = 33793
NUM_OBS_ISING = 3
NUM_OBS_ISING_CORR = 4
NUM_OBS_AT_CORR = 16
DISORDER_UNIFORM = 0
DISORDER_GAUSSIAN = 1
NUM_OBS_AT = 12
DENSITY_AXIS_BINS = 32
DENSITY_CELLS = 1024

SIGNATURES = {
    "rewl_last_error": ([], C.c_char_p),
    "rewl_device_count": ([C.POINTER(C.c_int)], C.c_int),
    "rewl_version": ([], C.c_char_p),
    "rewl_create": ([C.POINTER(SystemDesc), C.c_int32, C.POINTER(Params), C.POINTER(_P)], C.c_int),
    "rewl_destroy": ([_P], C.c_int),
    "rewl_adjust_to_range": ([_P], C.c_int),
    "rewl_sweeps": ([_P, C.c_int32], C.c_int),
    "rewl_exchange": ([_P], C.c_int),
    "rewl_check": ([_P], C.c_int),
    "rewl_run": ([_P, C.c_int64], C.c_int),
    "rewl_finalize": ([_P], C.c_int),
    "rewl_synchronize": ([_P], C.c_int),
    "rewl_num_walkers": ([_P, C.POINTER(C.c_int32)], C.c_int),
    "rewl_num_windows": ([_P, C.POINTER(C.c_int32)], C.c_int),
    "rewl_get_window_info": ([_P, C.c_int32, C.POINTER(WindowInfo)], C.c_int),
    "rewl_get_stats": ([_P, C.POINTER(Stats)], C.c_int),
    "rewl_get_logdos": ([_P, C.c_int32, C.POINTER(C.c_double), C.c_int32], C.c_int),
    "rewl_get_counts": ([_P, C.c_int32, C.POINTER(C.c_uint32), C.c_int32], C.c_int),
    "rewl_get_visits": ([_P, C.c_int32, C.POINTER(C.c_uint64), C.c_int32], C.c_int),
    "rewl_get_spins": ([_P, C.c_int32, C.POINTER(C.c_int8), C.c_int32], C.c_int),
    "rewl_get_energy": ([_P, C.c_int32, C.POINTER(C.c_double)], C.c_int),
    "rewl_get_flavor": ([_P, C.c_int32, C.POINTER(C.c_int32)], C.c_int),
    "rewl_get_rng_cursor": ([_P, C.c_int32, C.POINTER(C.c_uint64)], C.c_int),
    "rewl_get_observables": ([_P, C.c_int32, C.POINTER(C.c_double), C.c_int32, C.c_int32], C.c_int),
    "rewl_get_densities": ([_P, C.c_int32, C.POINTER(C.c_uint32), C.c_int32], C.c_int),
    "rewl_anneal": ([C.POINTER(SystemDesc), C.c_int32, C.POINTER(AnnealParams), C.POINTER(C.c_double),
                     C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.POINTER(C.c_int8)], C.c_int),
    "rewl_thermo_observables": ([C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32,
                                 C.c_int32, C.c_double, C.c_double, C.c_int32, C.POINTER(C.c_double)], C.c_int),
    "rewl_random_field": ([C.c_uint64, C.c_int32, C.c_double, C.c_int32, C.POINTER(C.c_double)], C.c_int),
    "rewl_set_logdos": ([_P, C.c_int32, C.POINTER(C.c_double), C.c_int32], C.c_int),
    "rewl_set_window_lnf": ([_P, C.c_int32, C.c_double], C.c_int),
    "rewl_set_spins": ([_P, C.c_int32, C.POINTER(C.c_int8), C.c_int32], C.c_int),
    "rewl_refine": ([_P, C.c_int64, C.c_int32], C.c_int),
    "rewl_refine_begin": ([_P], C.c_int),
    "rewl_refine_sweeps": ([_P, C.c_int64, C.c_int32], C.c_int),
    "rewl_refine_end": ([_P], C.c_int),
    "rewl_set_system_spins": ([_P, C.c_int32, C.POINTER(C.c_int8), C.c_int32], C.c_int),
    "rewl_stitch": ([_P, C.c_int32, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32,
                     C.POINTER(C.c_int32)], C.c_int),
    "rewl_stitch_overlap": ([_P, C.c_int32, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32,
                     C.POINTER(C.c_int32)], C.c_int),
    "rewl_thermodynamics": ([C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32, C.c_int32, C.c_double,
                             C.c_double, C.c_int32, C.POINTER(C.c_double)], C.c_int),
    "rewl_thermodynamics_noalloc": ([C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32, C.c_int32,
                                     C.c_double, C.c_double, C.c_int32, C.POINTER(C.c_double)], C.c_int),
    "rewl_selfcheck": ([C.c_int32, C.POINTER(C.c_double)], C.c_int),
    # include/rewl_mlp.h
    "rewl_mlp_create": ([C.c_uint64, C.POINTER(_P)], C.c_int),
    "rewl_mlp_destroy": ([_P], C.c_int),
    "rewl_mlp_get_params": ([_P, C.POINTER(C.c_float), C.c_int64], C.c_int),
    "rewl_mlp_set_params": ([_P, C.POINTER(C.c_float), C.c_int64], C.c_int),
    "rewl_mlp_forward": ([_P, C.POINTER(C.c_float), C.c_int64, C.POINTER(C.c_float)], C.c_int),
    "rewl_mlp_gradient": ([_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64,
                           C.POINTER(C.c_float), C.POINTER(C.c_double)], C.c_int),
    "rewl_mlp_train": ([_P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64,
                        C.POINTER(MlpTrainParams), C.POINTER(C.c_float)], C.c_int),
    "rewl_mlp_last_ms": ([_P, C.POINTER(C.c_double)], C.c_int),
}

_lib = None


def load_engine(path=ENGINE_SO):
    """Load librewl_hip.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"HIP engine not built: {path} is missing (run python -m rewl_simulator_amd.build)")
    lib = C.CDLL(path)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


class REWLError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check(code):
    if code != REWL_OK:
        raise REWLError(code, _lib.rewl_last_error().decode())
    return code
