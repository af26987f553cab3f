"""ctypes binding of libsimcogpu.so (the C ABI declared in include/simcogpu.h).

The library is loaded from the package's in-tree build directory.  There is
no fallback: if it is missing the import fails with the build command.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libsimcogpu.so")

SG_OK = 0
SG_NOT_CONVERGED = 1
SG_NAN = 2
SG_GUARD = 3
SG_ERR_ARG = -1
SG_ERR_HIP = -2
SG_ERR_ALLOC = -3
SG_ERR_STATE = -4
SG_PREC_FP64 = 0
SG_PREC_MIXED = 1

# enum sg_field
FIELDS = ["VOF", "PHI", "VOFS", "NYS", "WLLH", "NXS", "PHIS", "NX", "NY", "MRES", "VOFSU", "VOFSV",
          "U", "V", "P", "UPRED", "VPRED", "PP", "DPU", "DPV", "AE", "AW", "AN", "AS",
          "UFLAG", "VFLAG",
          "VOFO", "PHIO", "NXO", "NYO", "UVOF", "VVOF", "UVOFO", "VVOFO",
          "FLUXN1U", "FLUXN1V",
          "UPHI", "UNX", "UNY", "VPHI", "VNX", "VNY"]
F = {name: k for k, name in enumerate(FIELDS)}
CELL_FIELDS = set(FIELDS[:12]) | set(FIELDS[26:34]) | set(FIELDS[36:42])
GHOST_FIELDS = set(FIELDS[12:24]) | {"FLUXN1U", "FLUXN1V"}
INT_FIELDS = {"UFLAG", "VFLAG"}


class Params(ctypes.Structure):
    _fields_ = [
        ("rho_l", ctypes.c_double), ("rho_g", ctypes.c_double), ("rho_ref", ctypes.c_double),
        ("nu_l", ctypes.c_double), ("nu_g", ctypes.c_double), ("nu_ref", ctypes.c_double),
        ("Re", ctypes.c_double), ("Fr", ctypes.c_double),
        ("bc_w", ctypes.c_int), ("bc_e", ctypes.c_int), ("bc_s", ctypes.c_int), ("bc_n", ctypes.c_int),
        ("precision", ctypes.c_int), ("max_iter", ctypes.c_int), ("tol", ctypes.c_double),
        ("mg_pre", ctypes.c_int), ("mg_post", ctypes.c_int), ("warm_start", ctypes.c_int),
        ("check_every", ctypes.c_int), ("mg_coarse_krylov", ctypes.c_int), ("momentum_method", ctypes.c_int),
    ]


class SolveInfo(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int), ("status", ctypes.c_int), ("relres", ctypes.c_double),
        ("max_row_residual", ctypes.c_double), ("ms", ctypes.c_double), ("setup_ms", ctypes.c_double),
        ("warm_ms", ctypes.c_double), ("warm_relres", ctypes.c_double), ("warm_energy", ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


STAG_CELL = ["vofS", "nxS", "nyS", "phiS", "WlLh", "delh", "ccx", "ccy"]
STAG_INT = ["moex", "active"]
STAG_GHOST = ["aE", "aW", "aN", "aS", "EtaE", "EtaN", "AlE", "AlN", "DAlE", "DAlN", "SxE", "SyN",
              "fcex", "fcey", "fcnx", "fcny"]


class StagGeometry(ctypes.Structure):
    _fields_ = ([(k, ctypes.POINTER(ctypes.c_double)) for k in STAG_CELL] +
                [(k, ctypes.POINTER(ctypes.c_int)) for k in STAG_INT] +
                [(k, ctypes.POINTER(ctypes.c_double)) for k in STAG_GHOST])


class PredictorInfo(ctypes.Structure):
    _fields_ = [("iterations_u", ctypes.c_int), ("iterations_v", ctypes.c_int),
                ("relres_u", ctypes.c_double), ("relres_v", ctypes.c_double), ("ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ParticleParams(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("rho_l", "rho_g", "nu_l", "nu_g", "rho_p", "gx", "gy", "lref", "uref",
                                               "us", "vs")]


class ClsvofParams(ctypes.Structure):
    _fields_ = [("nsub", ctypes.c_int), ("depth", ctypes.c_double), ("amp", ctypes.c_double),
                ("kw", ctypes.c_double), ("omega", ctypes.c_double)]


# exported symbol -> (argtypes, restype); kept in step with include/simcogpu.h
# solid shapes of sg_geometry_rebuild
SOLID_KINDS = {"none": 0, "box": 1, "circle": 2, "ellipse": 3, "boom": 4}


class SolidShape(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("cx", ctypes.c_double), ("cy", ctypes.c_double), ("a", ctypes.c_double),
                ("b", ctypes.c_double), ("c", ctypes.c_double)]


# per-step diagnostics ring (SG_DIAG_COLS, SG_DIAG_RING); the columns
DIAG_COLS = 8
DIAG_RING = 4096
DIAG_NAMES = ("dt", "liquid_volume", "gas_volume", "gas_centroid_x", "gas_centroid_y", "max_abs_mres", "max_abs_u",
              "max_abs_v")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_V = ctypes.c_void_p
_CTX = ctypes.c_void_p
_F = ctypes.POINTER(ctypes.c_float)
PROTOTYPES = {
    "sg_last_error": ([], ctypes.c_char_p),
    "sg_params_default": ([ctypes.POINTER(Params)], ctypes.c_int),
    "sg_ctx_create": ([ctypes.c_int, ctypes.c_int, _D, _D, _D, _D, ctypes.POINTER(Params),
                       ctypes.POINTER(_CTX)], ctypes.c_int),
    "sg_ctx_destroy": ([_CTX], ctypes.c_int),
    "sg_set_params": ([_CTX, ctypes.POINTER(Params)], ctypes.c_int),
    "sg_field_ptr": ([_CTX, ctypes.c_int], ctypes.c_void_p),
    "sg_upload": ([_CTX, ctypes.c_int, _V], ctypes.c_int),
    "sg_download": ([_CTX, ctypes.c_int, _V], ctypes.c_int),
    "sg_synchronize": ([_CTX], ctypes.c_int),
    "sg_set_geometry": ([_CTX, _D, _D, _D, _D, _D, _D, _D, _I, _I, _D, _D], ctypes.c_int),
    "sg_poisson_build_resident": ([_CTX, ctypes.c_double], ctypes.c_int),
    "sg_poisson_solve_resident": ([_CTX, ctypes.POINTER(SolveInfo)], ctypes.c_int),
    "sg_poisson_solve": ([_CTX, _D, _D, _D, _D, _D, _D, ctypes.c_double, _D, ctypes.POINTER(SolveInfo)],
                         ctypes.c_int),
    "sg_poisson_matrix": ([_CTX, _D], ctypes.c_int),
    "sg_poisson_rhs": ([_CTX, _D], ctypes.c_int),
    "sg_poisson_apply": ([_CTX, _D, _D], ctypes.c_int),
    "sg_velocity_correct_resident": ([_CTX, ctypes.c_double], ctypes.c_int),
    "sg_clsvof_params_default": ([ctypes.POINTER(ClsvofParams)], ctypes.c_int),
    "sg_set_centres": ([_CTX, _D, _D], ctypes.c_int),
    "sg_clsvof_advect_resident": ([_CTX, ctypes.c_double, ctypes.c_double, ctypes.POINTER(ClsvofParams)],
                                  ctypes.c_int),
    "sg_clsvof_advect": ([_CTX, _D, _D, _D, _D, _D, _D, ctypes.c_double, ctypes.c_double,
                          ctypes.POINTER(ClsvofParams)], ctypes.c_int),
    "sg_time_kernel": ([_CTX, ctypes.c_int, ctypes.c_int, _D], ctypes.c_int),
    "sg_set_staggered_geometry": ([_CTX, ctypes.c_int, ctypes.POINTER(StagGeometry)], ctypes.c_int),
    "sg_predictor_resident": ([_CTX, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                               ctypes.POINTER(PredictorInfo)], ctypes.c_int),
    "sg_predictor_rows": ([_CTX, _D, _D, _D, _D], ctypes.c_int),
    "sg_interface_to_uv_resident": ([_CTX], ctypes.c_int),
    "sg_particles_set": ([_CTX, ctypes.c_int, _D, _D, _D, _D, _D], ctypes.c_int),
    "sg_particles_get": ([_CTX, ctypes.c_int, _D, _D, _D, _D], ctypes.c_int),
    "sg_particles_track": ([_CTX, ctypes.c_double, ctypes.POINTER(ParticleParams), _I], ctypes.c_int),
    "sg_geometry_rebuild": ([_CTX, ctypes.POINTER(SolidShape)], ctypes.c_int),
    "sg_stag_download": ([_CTX, ctypes.c_int, ctypes.c_int, _V], ctypes.c_int),
    "sg_object_force": ([_CTX, ctypes.c_double, _D, _D, _D], ctypes.c_int),
    "sg_set_moving_solid": ([_CTX, ctypes.POINTER(SolidShape), ctypes.c_double, ctypes.c_double, ctypes.c_double,
                             ctypes.c_double], ctypes.c_int),
    "sg_get_moving_solid": ([_CTX, ctypes.POINTER(SolidShape), _D, _D], ctypes.c_int),
    "sg_geometry_snapshot": ([_CTX], ctypes.c_int),
    "sg_inter_new_var": ([_CTX, ctypes.c_double], ctypes.c_int),
    "sg_diagnostics_record": ([_CTX], ctypes.c_int),
    "sg_diagnostics_read": ([_CTX, ctypes.c_int, _D, _I, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
    "sg_state_bytes": ([_CTX, ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
    "sg_state_save": ([_CTX, _V, ctypes.c_size_t], ctypes.c_int),
    "sg_state_load": ([_CTX, _V, ctypes.c_size_t], ctypes.c_int),
    "sg_upload_device": ([_CTX, ctypes.c_int, _V], ctypes.c_int),
    "sg_time_step_size": ([_CTX, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                           ctypes.c_double, _D], ctypes.c_int),
    "sg_time_step_size_noalloc": ([ctypes.c_int, ctypes.c_int, _V, _V, _V, _V] + [ctypes.c_double] * 6 + [_V, _V],
                                  ctypes.c_int),
    "sg_closed_box_bc": ([_CTX], ctypes.c_int),
    "sg_field_copy": ([_CTX, ctypes.c_int, ctypes.c_int], ctypes.c_int),
    "sg_download_device": ([_CTX, ctypes.c_int, _V], ctypes.c_int),
    "sg_nn_num_weights": ([], ctypes.c_size_t),
    "sg_nn_init_weights": ([_F, ctypes.c_ulonglong], ctypes.c_int),
    "sg_nn_create": ([ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
    "sg_nn_destroy": ([_V], ctypes.c_int),
    "sg_nn_set_weights": ([_V, _F], ctypes.c_int),
    "sg_nn_get_weights": ([_V, _F], ctypes.c_int),
    "sg_nn_forward": ([_V, _V, _V, ctypes.c_int], ctypes.c_int),
    "sg_nn_gradient": ([_V, _V, _V, _V, ctypes.c_int, _F, _F], ctypes.c_int),
    "sg_nn_train_step": ([_V, _V, _V, _V, ctypes.c_int, ctypes.c_float, _F], ctypes.c_int),
    "sg_nn_level_dims": ([_CTX, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
                         ctypes.c_int),
    "sg_nn_features": ([_CTX, ctypes.c_int, _V, _V, _V, _D], ctypes.c_int),
    "sg_set_learned": ([_CTX, _F, ctypes.c_int, ctypes.c_double, ctypes.c_int], ctypes.c_int),
    "sg_set_learned_prediction": ([_CTX, ctypes.c_int, _D, ctypes.c_int], ctypes.c_int),
    "sg_sweep_timing": ([_CTX, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong), _D], ctypes.c_int),
}

_lib = None


class SimcoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"simcogpu error {code}: {msg}")
        self.code = code


def load():
    """Load libsimcogpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m simcoflow_amd.build`")
    # torch carries its own HIP runtime beside the system one this library
    # links; two runtimes share a process only if torch's starts first (one
    # started after the library's sees no GPU), so start it before any
    # library call when torch is installed
    try:
        import torch
        torch.cuda.is_available()
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    for name, (argt, rest) in PROTOTYPES.items():
        fn = getattr(L, name)
        fn.argtypes = argt
        fn.restype = rest
    _lib = L
    return L


def last_error():
    return load().sg_last_error().decode(errors="replace")


def check(rc, allow=()):
    if rc != SG_OK and rc not in allow:
        raise SimcoError(rc, last_error())
    return rc
