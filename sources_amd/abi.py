"""ctypes mirror of include/sphg.h and loader of the HIP library.

The structures here are byte-for-byte the C structures of the public ABI; the
Fortran host (sources_amd/fortran/sphg_mod.f90) declares the same layout with
``bind(C)`` derived types.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsphg.so")

SPHG_OK = 0
SPHG_ERR_TOO_MANY_INTERACTIONS = -1
SPHG_ERR_OUT_OF_RANGE = -2
SPHG_ERR_BAD_KERNEL = -3
SPHG_ERR_INVALID = -4
SPHG_ERR_HIP = -5
SPHG_ERR_UNSUPPORTED = -6
SPHG_ERR_NONFINITE = -7
SPHG_ERR_OUT_OF_MEMORY = -8
SPHG_ERR_MOVED_TOO_FAR = -9

SPHG_WATER = 1
SPHG_SOIL = 2
SPHG_FP64 = 0
SPHG_MIXED = 1

F_X = 1 << 0
F_VX = 1 << 1
F_RHO = 1 << 2
F_P = 1 << 3
F_C = 1 << 4
F_VOF = 1 << 5
F_STRESS = 1 << 6
F_EPSP = 1 << 7
F_DRHO = 1 << 8
F_DVX = 1 << 9
F_DVOF = 1 << 10
F_DSTRESS = 1 << 11
F_AV = 1 << 12
F_STRAIN = 1 << 13
F_VOF2 = 1 << 14
F_FAIL = 1 << 15
F_COUNTIAC = 1 << 16
F_MIN = 1 << 17
F_ALL = (1 << 64) - 1

ERROR_TEXT = {
    SPHG_OK: "ok",
    SPHG_ERR_TOO_MANY_INTERACTIONS: ">>> ERROR <<< : Too many interactions",
    SPHG_ERR_OUT_OF_RANGE: ">>> ERROR <<< : Particle out of range",
    SPHG_ERR_BAD_KERNEL: "get_scale_k: No such kernel function!",
    SPHG_ERR_INVALID: "invalid argument",
    SPHG_ERR_HIP: "HIP runtime error",
    SPHG_ERR_UNSUPPORTED: "switch combination not supported",
    SPHG_ERR_NONFINITE: "non-finite value detected",
    SPHG_ERR_OUT_OF_MEMORY: "out of device memory",
    SPHG_ERR_MOVED_TOO_FAR: "a particle moved farther than the neighbour-list skin allows",
}


class SphgError(RuntimeError):
    """Raised for a non-zero return code; `code` carries the ABI code."""

    def __init__(self, code, where=""):
        self.code = code
        super().__init__(f"{where}: {ERROR_TEXT.get(code, 'error')} (code {code})")


class Params(C.Structure):
    _fields_ = [
        ("water_rho0", C.c_double), ("water_b", C.c_double), ("water_gamma", C.c_double),
        ("water_c", C.c_double), ("water_viscosity", C.c_double),
        ("soil_rho0", C.c_double), ("soil_k", C.c_double), ("soil_E", C.c_double),
        ("soil_niu", C.c_double), ("soil_c", C.c_double),
        ("soil_porosity", C.c_double), ("soil_permeability", C.c_double),
        ("soil_cohesion", C.c_double), ("soil_phi", C.c_double),
        ("nnps", C.c_int32), ("skf", C.c_int32),
        ("numeric_gravity", C.c_double),
        ("alpha", C.c_double), ("beta", C.c_double), ("etq", C.c_double),
        ("dd", C.c_double), ("p1", C.c_double), ("p2", C.c_double),
        ("delta", C.c_double), ("epsilon", C.c_double), ("gravity", C.c_double),
        ("single_phase", C.c_int32), ("summation_density", C.c_int32),
        ("average_velocity", C.c_int32), ("visc_artificial", C.c_int32),
        ("artificial_density", C.c_int32), ("soil_artificial_stress", C.c_int32),
        ("volume_fraction", C.c_int32), ("water_artificial_volume", C.c_int32),
        ("volume_fraction_renorm", C.c_int32), ("stress_integration", C.c_int32),
        ("yield_criterion", C.c_int32), ("plasticity", C.c_int32),
        ("x_mingeom", C.c_double), ("x_maxgeom", C.c_double),
        ("y_mingeom", C.c_double), ("y_maxgeom", C.c_double),
        ("maxngx", C.c_int32), ("maxngy", C.c_int32),
        ("dt", C.c_double),
        ("precision", C.c_int32),
        ("integrate_scheme", C.c_int32),
        ("reserved", C.c_int32 * 6),
    ]


_PI = C.POINTER(C.c_int32)
_PD = C.POINTER(C.c_double)

VIEW_INT_ARRAYS = ("itype", "fail", "countiac", "pair_i", "pair_j")
VIEW_DBL_ARRAYS = (
    "x", "vx", "mass", "hsml",
    "rho", "p", "c", "vof", "vof2", "dvof", "vof_min",
    "rho_min", "v_min", "drho", "dvx", "av",
    "sxx", "syy", "sxy", "txx", "tyy", "txy", "wxy", "vcc", "epsilon_p",
    "dsxx", "dsyy", "dsxy", "dp",
    "w", "dwdx",
    "sxx_min", "sxy_min", "syy_min", "p_min",
    "dsxx2", "dsxy2", "dsyy2", "dp2",
)


class PhaseView(C.Structure):
    _fields_ = (
        [("ntotal", C.c_int32), ("nvirt", C.c_int32), ("maxn", C.c_int32),
         ("max_interaction", C.c_int32), ("material", C.c_int32),
         ("niac", C.c_int32), ("nfail", C.c_int32), ("reserved0", C.c_int32),
         ("dspp", C.c_double)]
        + [(n, _PI) for n in VIEW_INT_ARRAYS]
        + [(n, _PD) for n in VIEW_DBL_ARRAYS]
    )


_lib = None


def load_library(path=None):
    """Load libsphg.so.  Raises if it is not built: there is no fallback path."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    # RTLD_DEEPBIND: the library's HIP calls bind to the HIP runtime it links
    # (ROCm's libamdhip64.so.7).  Without it, a process that imported torch
    # first has torch's own bundled libamdhip64.so in the global symbol scope,
    # and every HIP symbol of the library (kernel registration, streams,
    # graph capture) would bind to that other runtime instead.
    lib = C.CDLL(p, mode=os.RTLD_LAZY | os.RTLD_DEEPBIND)
    vp = C.c_void_p
    lib.sphg_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    lib.sphg_destroy.argtypes = [vp]
    lib.sphg_destroy.restype = None
    lib.sphg_set_params.argtypes = [vp, C.POINTER(Params)]
    lib.sphg_bind_phase.argtypes = [vp, C.c_int, C.POINTER(PhaseView)]
    lib.sphg_sync.argtypes = [vp, C.c_int, C.c_uint64, C.c_int]
    lib.sphg_set_itimestep.argtypes = [vp, C.c_int]
    lib.sphg_get_itimestep.argtypes = [vp]
    for name in ("sphg_direct_find", "sphg_link_list", "sphg_con_density", "sphg_sum_density",
                 "sphg_delta_sph_density", "sphg_internal_force", "sphg_art_visc"):
        getattr(lib, name).argtypes = [vp, C.c_int]
    lib.sphg_direct_find_2.argtypes = [vp]
    lib.sphg_single_step.argtypes = [vp, C.c_int]
    lib.sphg_time_integration.argtypes = [vp, C.c_int, C.c_int]
    lib.sphg_strerror.argtypes = [C.c_int]
    lib.sphg_strerror.restype = C.c_char_p
    lib.sphg_last_hip_error.argtypes = [vp]
    lib.sphg_stream.argtypes = [vp]
    lib.sphg_stream.restype = vp
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int)
    lib.sphg_gnn_nparams.argtypes = []
    lib.sphg_gnn_create.argtypes = [vp, C.c_int, C.c_int, C.c_uint64, C.POINTER(vp)]
    lib.sphg_gnn_destroy.argtypes = [vp]
    lib.sphg_gnn_destroy.restype = None
    lib.sphg_gnn_weights.argtypes = [vp, fp, C.c_int]
    lib.sphg_gnn_num_samples.argtypes = [vp]
    lib.sphg_gnn_record.argtypes = [vp]
    lib.sphg_gnn_train.argtypes = [vp, C.c_int, C.c_float, fp]
    lib.sphg_gnn_step.argtypes = [vp, C.c_int]
    lib.sphg_gnn_rollout.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_int, ip, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]
    lib.sphg_gnn_sample_size.argtypes = [vp, C.c_int]
    lib.sphg_gnn_debug.argtypes = [vp, C.c_int, fp, fp, ip, ip, fp, fp, fp, fp, fp, fp]
    lib.sphg_gnn_gains.argtypes = [vp, fp]
    lib.sphg_batch_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(vp)]
    lib.sphg_batch_destroy.argtypes = [vp]
    lib.sphg_batch_destroy.restype = None
    lib.sphg_batch_run.argtypes = [vp, C.c_int]
    lib.sphg_batch_set_streams.argtypes = [vp, C.c_int]
    lib.sphg_batch_tune.argtypes = [vp, C.c_int, ip, C.POINTER(C.c_double)]
    lib.sphg_batch_staged.argtypes = [vp]
    lib.sphg_profile_enable.argtypes = [vp, C.c_int]
    lib.sphg_profile_reset.argtypes = [vp]
    lib.sphg_profile_get.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    lib.sphg_pair_counts.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
    lib.sphg_nlist_stats.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    # device pointers are passed as integers (c_void_p)
    lib.sphg_direct_find_noalloc.argtypes = [C.POINTER(Params), C.c_int, vp, vp, C.c_int, vp, vp, vp, vp, vp,
                                             C.POINTER(C.c_int32), vp]
    if path is None:
        _lib = lib
    return lib


# Every symbol the public header declares; tests check the library exports them.
EXPORTED_SYMBOLS = (
    "sphg_create", "sphg_destroy", "sphg_set_params", "sphg_bind_phase", "sphg_sync",
    "sphg_set_itimestep", "sphg_get_itimestep",
    "sphg_direct_find", "sphg_link_list", "sphg_direct_find_2",
    "sphg_con_density", "sphg_sum_density", "sphg_delta_sph_density",
    "sphg_internal_force", "sphg_art_visc",
    "sphg_single_step", "sphg_time_integration",
    "sphg_strerror", "sphg_last_hip_error", "sphg_stream",
    "sphg_profile_enable", "sphg_profile_reset", "sphg_profile_get", "sphg_pair_counts",
    "sphg_nlist_stats",
    "sphg_gnn_nparams", "sphg_gnn_create", "sphg_gnn_destroy", "sphg_gnn_weights", "sphg_gnn_num_samples",
    "sphg_gnn_record", "sphg_gnn_train", "sphg_gnn_step", "sphg_gnn_rollout", "sphg_gnn_sample_size",
    "sphg_gnn_debug", "sphg_gnn_gains",
    "sphg_batch_create", "sphg_batch_destroy", "sphg_batch_run", "sphg_batch_set_streams", "sphg_batch_tune",
    "sphg_batch_staged",
    "sphg_direct_find_noalloc",
)

# Limits of sphg_direct_find_noalloc (include/sphg.h)
NOALLOC_MAX_N = 131072
NOALLOC_MAX_CELLS = 262144
