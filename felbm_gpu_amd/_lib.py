"""ctypes binding of ``include/felbm_mi355x.h`` (libfelbm_mi355x.so).

Loading fails loudly when the library has not been built: there is no
fallback path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfelbm_mi355x.so")

FELBM_OK = 0
FELBM_ENONFINITE = -6
ERRORS = {-1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "ESTATE", -5: "EUNSUPPORTED", -6: "ENONFINITE"}


class FelbmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "E?"), code, msg))
        self.code = code


class Geometry(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("periodic", C.c_int * 3),
                ("use_halfway_bb", C.c_int), ("solid", C.POINTER(C.c_uint8)),
                ("site_xyz", C.POINTER(C.c_int32)), ("n_sites", C.c_int64)]


class Params(C.Structure):
    _fields_ = [("rho0", C.c_double), ("rho1", C.c_double), ("kappa", C.c_double), ("beta", C.c_double),
                ("boundary_coefficient", C.c_double), ("tau0", C.c_double), ("tau1", C.c_double),
                ("mobility", C.c_double), ("forcing_factor", C.c_double),
                ("acceleration", C.c_double * 3), ("forcing", C.c_double * 3), ("mrt_lambda", C.c_double)]


class Switches(C.Structure):
    _fields_ = [("use_mrt", C.c_int), ("correct_op_mass", C.c_int), ("stream_matrix_free", C.c_int),
                ("grad_matrix_free", C.c_int), ("fused", C.c_int), ("fuse_collision", C.c_int),
                ("mrt_fast_transform", C.c_int), ("stream_inplace", C.c_int)]


class OpenBnd(C.Structure):
    _fields_ = [("use_in_vel", C.c_int), ("use_in_prs", C.c_int), ("use_in_fluid", C.c_int),
                ("use_out_prs", C.c_int), ("use_out_fluid", C.c_int), ("u_in", C.c_double * 3),
                ("p_in", C.c_double), ("p_out", C.c_double), ("c_out_fixed", C.c_double),
                ("rho_out_fixed", C.c_double), ("inlet_c_a", C.c_double), ("inlet_mode", C.c_int),
                ("split_axis", C.c_int), ("inlet_period", C.c_double), ("inlet_duty", C.c_double),
                ("inlet_ramp", C.c_double), ("split_pos", C.c_double)]


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_PD = C.POINTER(C.c_double)
_PI32 = C.POINTER(C.c_int32)

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "felbm_create": (_I, [C.POINTER(_P), _I, _I]),
    "felbm_destroy": (_I, [_P]),
    "felbm_add_instance": (_I, [_P, C.POINTER(Geometry), C.POINTER(Params), C.POINTER(Switches), C.POINTER(_I)]),
    "felbm_remove_instance": (_I, [_P, _I]),
    "felbm_num_sites": (_I, [_P, _I, C.POINTER(C.c_int64)]),
    "felbm_state_size": (_I, [_P, _I, C.POINTER(C.c_int64)]),
    "felbm_num_instances": (_I, [_P, C.POINTER(_I)]),
    "felbm_upload_state": (_I, [_P, _I, _PD, _PD]),
    "felbm_download_state": (_I, [_P, _I, _PD, _PD]),
    "felbm_init_rest": (_I, [_P, _I, _PD, _D]),
    "felbm_step": (_I, [_P, _I, _I]),
    "felbm_step_graph": (_I, [_P, _I, _I]),
    "felbm_set_forcing_factor": (_I, [_P, _I, _D]),
    "felbm_forcing_factor_at": (_I, [C.c_char_p, _D, C.c_int64, C.POINTER(C.c_double)]),
    "felbm_tracers_init": (_I, [_P, _I, C.c_int64, _PD, _PD, C.c_uint64, _I, _I]),
    "felbm_tracers_download": (_I, [_P, _I, _PD, _PD, _PI32]),
    "felbm_tracers_count": (_I, [_P, _I, C.POINTER(C.c_int64)]),
    "felbm_tracers_set_step": (_I, [_P, _I, C.c_uint64]),
    "felbm_tracers_stretching": (_I, [_P, _I, _I, _PD, _PD, C.c_uint64, _I, _I]),
    "felbm_tracers_stretch_rows": (_I, [_P, _I, _PD, _I, C.POINTER(C.c_int), C.POINTER(C.c_int64)]),
    "felbm_tracers_stretch_state": (_I, [_P, _I, _PD, _PD]),
    "felbm_synchronize": (_I, [_P]),
    "felbm_record_target_mass": (_I, [_P, _I]),
    "felbm_get_target_mass": (_I, [_P, _I, _PD]),
    "felbm_set_target_mass": (_I, [_P, _I, _D]),
    "felbm_current_mass": (_I, [_P, _I, _PD]),
    "felbm_flow_stats": (_I, [_P, _I, _PD]),
    "felbm_flow_stats_launch": (_I, [_P]),
    "felbm_flow_stats_ready": (_I, [_P, C.POINTER(_I)]),
    "felbm_flow_stats_collect": (_I, [_P, _I, _PD]),
    "felbm_download_fields": (_I, [_P, _I, _PD]),
    "felbm_set_open_bnd": (_I, [_P, _I, C.POINTER(OpenBnd), _PI32, _I, _PI32, _I, _PD, _PI32, _PI32, _I, _I]),
    "felbm_export_tables": (_I, [_P, _I] + [_PI32] * 10),
    "felbm_export_tables_noalloc": (_I, [_P, _I] + [_PI32] * 10),
    "felbm_mrt_matrices": (_I, [_PD, _PD]),
    "felbm_kernel_times": (_I, [_P, _I, C.POINTER(C.c_float)]),
    "felbm_autotune": (_I, [_P, _I, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    "felbm_set_block_sizes": (_I, [_P, C.POINTER(C.c_int)]),
    "felbm_copy_probe": (_I, [_I, C.c_int64, _I, _PD]),
    "felbm_stream_probe": (_I, [_I, C.c_int64, _I, _I, _PD]),
    "felbm_last_error": (C.c_char_p, []),
    "felbm_version": (C.c_char_p, []),
    "felbm_geom_cylinders": (_I, [_I, _I, _I, _I, _D, _D, _D, C.c_uint64, _I, C.POINTER(C.c_uint8)]),
    "felbm_geom_spheres": (_I, [_I, _I, _I, _D, _D, _D, C.c_uint64, C.POINTER(C.c_uint8)]),
    "felbm_keep_largest_cluster": (C.c_int64, [_I, _I, _I, _I, _I, _I, C.POINTER(C.c_uint8)]),
}

_lib = None


def load():
    """Load the engine library (built in-tree); raise if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libfelbm_mi355x.so is not built; run __graft_entry__.build() "
                          "(python -m felbm_gpu_amd.build)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code):
    if code < 0:
        raise FelbmError(code, load().felbm_last_error().decode(errors="replace"))
    return code
