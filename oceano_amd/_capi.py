"""ctypes binding of liboceano_dg.so (declarations in include/oceano_dg.h).

The library is built in-tree by oceano_amd/build.py.  Loading fails loudly when it is
missing: there is no fallback implementation of any entry point.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboceano_dg.so")

_lib = None

OC_OK = 0
OC_ERR_ARGUMENT = 1
OC_ERR_MESH = 2
OC_ERR_BOUNDARY = 3
OC_ERR_HIP = 4
OC_ERR_NONFINITE = 5
OC_ERR_UNSUPPORTED = 6
OC_NOALLOC_MAX_BATCHES = 16384

OC_BC_WALL = 1
OC_BC_SUPERCRITICAL_INFLOW = 2
OC_BC_SUPERCRITICAL_OUTFLOW = 3
OC_BC_HEIGHT_INFLOW = 4
OC_BC_DISCHARGE_INFLOW = 5
OC_BC_ABSORBING_OUTFLOW = 6

OC_PRECISION_FP64 = 0
OC_PRECISION_MIXED = 1

OC_STORE_CELL_GAUSS = 0
OC_STORE_CELL_LOBATTO = 1
OC_STORE_CELL_COLLOC = 2
OC_STORE_FACE = 3
OC_STORE_BOUNDARY = 4

CELL_BATCH = 64

P = C.c_void_p
I32 = C.c_int32
I64 = C.c_int64
U32 = C.c_uint32
U64 = C.c_uint64
D = C.c_double
PD = C.POINTER(C.c_double)
PF = C.POINTER(C.c_float)
PI32 = C.POINTER(C.c_int32)
PI64 = C.POINTER(C.c_int64)


class MeshDesc(C.Structure):
    _fields_ = [("n_cells", I32), ("cell_geom", PD), ("n_inner", I32), ("inner", PI32),
                ("n_boundary", I32), ("boundary", PI32)]


class Params(C.Structure):
    _fields_ = [("degree", I32), ("g", D), ("threshold_wetdry", D), ("viscosity", D),
                ("max_iteration_height", I32), ("precision", I32),
                ("n_boundary_kinds", I32), ("boundary_ids", PI32), ("boundary_kinds", PI32)]


# exported symbol -> (restype, argtypes)
SIGNATURES = {
    "oc_last_error": (C.c_char_p, []),
    "oc_set_option": (C.c_int, [C.c_char_p, C.c_int64]),
    "oc_dg_velocity": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "oc_dg_transport_speed": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.POINTER(C.c_double), C.c_void_p]),
    "oc_dg_volume": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.c_void_p]),
    "oc_dg_sum": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_double),
                            C.c_void_p]),
    "oc_rk_create": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double),
                               C.POINTER(C.c_void_p)]),
    "oc_rk_destroy": (None, [C.c_void_p]),
    "oc_rk_step": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                             C.c_void_p]),
    "oc_rk_capture": (C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_void_p]),
    "oc_rk_replay": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "oc_rk_create_multi": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p), C.c_int32,
                                     C.POINTER(C.c_double), C.POINTER(C.c_void_p)]),
    "oc_rk_step_multi": (C.c_int, [C.c_void_p, C.POINTER(C.c_double),
                                   C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                   C.c_void_p]),
    "oc_rk_capture_multi": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double),
                                      C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                      C.c_void_p]),
    "oc_rk_capture_steps": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p]),
    "oc_rk_set_time": (C.c_int, [C.c_void_p, C.c_double]),
    "oc_rk_time": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "oc_rk_registers": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                  C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "oc_mesh_create_rectangle": (C.c_int, [D, D, D, D, I32, I32, I32, I32, PI32, C.POINTER(P)]),
    "oc_mesh_read_msh": (C.c_int, [C.c_char_p, I32, I32, C.POINTER(P)]),
    "oc_mesh_refine_global": (C.c_int, [P, I32]),
    "oc_mesh_refine_fraction": (C.c_int, [P, D, U64]),
    "oc_mesh_info": (C.c_int, [P, PI64]),
    "oc_mesh_export": (C.c_int, [P, PD, PI32, PI32, PI32]),
    "oc_mesh_destroy": (None, [P]),
    "oc_dg_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(Params), C.POINTER(P)]),
    "oc_dg_destroy": (C.c_int, [P]),
    "oc_dg_n_points": (C.c_int, [P, I32, PI64]),
    "oc_dg_quadrature_points": (C.c_int, [P, I32, PD]),
    "oc_dg_set_static_data": (C.c_int, [P, I32, PD]),
    "oc_dg_set_boundary_values": (C.c_int, [P, PD]),
    "oc_dg_set_boundary_table": (C.c_int, [P, I32, PD, PD]),
    "oc_dg_sizes": (C.c_int, [P, PI64]),
    "oc_dg_vector_length": (C.c_int, [P, I32, PI64]),
    "oc_dg_pack": (C.c_int, [P, I32, P, P, P]),
    "oc_dg_unpack": (C.c_int, [P, I32, P, P, P]),
    "oc_dg_perform_stage_hydro": (C.c_int, [P, U32, D, PD, U32, P, P, P, C.POINTER(P),
                                            C.POINTER(P), U32, P, P, P, P, I32, P]),
    "oc_dg_perform_stage_hydro_multi": (C.c_int, [I32, C.POINTER(P), U32, PD,
                                                  C.POINTER(PD), U32,
                                                  C.POINTER(P), C.POINTER(P),
                                                  C.POINTER(C.POINTER(P)),
                                                  C.POINTER(C.POINTER(P)), U32, C.POINTER(P),
                                                  C.POINTER(P), C.POINTER(P), C.POINTER(P),
                                                  P]),
    "oc_dg_perform_stage_hydro_noalloc": (C.c_int, [P, U32, D, PD, U32, P, P, P,
                                                    C.POINTER(P), C.POINTER(P), U32, P, P, P,
                                                    P, I32, P]),
    "oc_dg_perform_stage_hydro_ark": (C.c_int, [P, U32, D, PD, PD, U32, D, P, P, P,
                                                C.POINTER(P), U32, C.POINTER(P), U32, P, P,
                                                P, P, P]),
    "oc_dg_apply": (C.c_int, [P, D, P, P, P, P, P, P]),
    "oc_dg_rhs": (C.c_int, [P, D, P, P, P, P, P, P]),
    "oc_dg_count_nonfinite": (C.c_int, [P, P, I64, PI64, P]),
    "oc_dg_last_sweeps": (C.c_int, [P, PI64, P]),
    "oc_dg_project_collocation": (C.c_int, [P, I32, PD, PD]),
    "oc_dg_compute_errors": (C.c_int, [P, PD, PD, PD, PD]),
    "oc_dg_set_timing": (C.c_int, [P, I32]),
    "oc_dg_timing_report": (C.c_int, [P, PD, PI64]),
    "oc_dg_set_subgrid_stats": (C.c_int, [P, PD, PI32]),
    "oc_dg_set_closure": (C.c_int, [P, P]),
    "oc_dg_closure_features": (C.c_int, [P, P, P, P, P]),
    "oc_dg_closure_source": (C.c_int, [P, P, P, P, P]),
    "oc_mlp_create": (C.c_int, [I32, I32, U64, C.POINTER(P)]),
    "oc_mlp_destroy": (None, [P]),
    "oc_mlp_n_params": (I32, []),
    "oc_mlp_dims": (C.c_int, [P, PI32, PI32]),
    "oc_mlp_get_params": (C.c_int, [P, PF]),
    "oc_mlp_set_params": (C.c_int, [P, PF]),
    "oc_mlp_set_normalization": (C.c_int, [P, PF, PF, PF, PF]),
    "oc_mlp_set_output_range": (C.c_int, [P, PF, PF]),
    "oc_mlp_forward": (C.c_int, [P, I64, P, P, P]),
    "oc_mlp_gradients": (C.c_int, [P, I64, P, P, PF, PD, P]),
    "oc_mlp_adam": (C.c_int, [P, C.c_float, C.c_float, C.c_float, C.c_float, P]),
    "oc_mlp_train_step": (C.c_int, [P, I64, P, P, C.c_float, PD, P]),
    "oc_icbc_vortex": (C.c_int, [PD, D, I64, PD, PD]),
    "oc_icbc_channel": (C.c_int, [PD, U64, I64, PD, PD]),
}


def lib():
    """Load liboceano_dg.so (once) and declare every exported function."""
    global _lib
    if _lib is None:
        # torch's HIP runtime first: loading the library before torch brings in a second
        # runtime, and the process then sees no device
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "oceano_amd: %s is missing; build it with `python oceano_amd/build.py` "
                "(there is no fallback path)" % LIB_PATH)
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class OceanoError(RuntimeError):
    def __init__(self, code, where):
        msg = lib().oc_last_error().decode(errors="replace")
        super().__init__("%s failed with status %d: %s" % (where, code, msg))
        self.code = code


def check(code, where):
    if code != OC_OK:
        raise OceanoError(code, where)


def dptr(a):
    """ctypes double* of a C-contiguous float64 numpy array."""
    return a.ctypes.data_as(PD)


def fptr(a):
    """ctypes float* of a C-contiguous float32 numpy array."""
    return a.ctypes.data_as(PF)


def iptr(a):
    return a.ctypes.data_as(PI32)
