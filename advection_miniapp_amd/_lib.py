"""ctypes binding of libdgops.so (the C ABI declared in include/dgops.h)."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdgops.so")

DG_OK = 0
DG_ERR_ARGUMENT = 1
DG_ERR_UNSUPPORTED = 2
DG_ERR_HIP = 3
DG_ERR_NO_CONVERGENCE = 4
DG_ERR_STATE = 5

DG_FP64, DG_FP32 = 0, 1
DG_OP_ADVECTION, DG_OP_ADVECTION_DIFFUSION, DG_OP_BURGERS = 0, 1, 2
DG_VEL_CONSTANT, DG_VEL_CELLULAR_3D, DG_VEL_VORTEX, DG_VEL_VORTEX_PERT = range(4)
(DG_SOL_ZERO, DG_SOL_WAVE, DG_SOL_SINE_PRODUCT, DG_SOL_GAUSSIAN,
 DG_SOL_POLYNOMIAL, DG_SOL_TWO_PULSES, DG_SOL_BURGERS_SINE) = range(7)
DG_LSRK_3_3, DG_LSRK_5_4, DG_LSRK_7_4, DG_LSRK_9_5 = range(4)
# steps per captured graph of dg_lsrk_time_steps
DG_GRAPH_STEPS = 8
DG_TIME_FACTOR_ONE, DG_TIME_FACTOR_COS = 0, 1


class DgConfig(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("degree", ctypes.c_int),
        ("n_components", ctypes.c_int),
        ("operator_kind", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("n_coarse", ctypes.c_int * 3),
        ("refine", ctypes.c_int),
        ("lo", ctypes.c_double * 3),
        ("hi", ctypes.c_double * 3),
        ("periodic", ctypes.c_int * 3),
        ("perturbation", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("velocity_kind", ctypes.c_int),
        ("velocity", ctypes.c_double * 3),
        ("time_factor", ctypes.c_int),
        ("final_time", ctypes.c_double),
        ("flux_alpha", ctypes.c_double),
        ("factor_skew", ctypes.c_double),
        ("viscosity", ctypes.c_double),
        ("penalty_factor", ctypes.c_double),
        ("solution_kind", ctypes.c_int),
        ("solution_param", ctypes.c_double * 8),
        ("irk_stages", ctypes.c_int),
        ("use_gl_quad", ctypes.c_int),
        ("use_gl_quad_mass", ctypes.c_int),
    ]


class DgInfo(ctypes.Structure):
    _fields_ = [
        ("n_cells", ctypes.c_int64),
        ("n_dofs", ctypes.c_int64),
        ("vector_size", ctypes.c_int64),
        ("dofs_per_cell", ctypes.c_int),
        ("cells_per_direction", ctypes.c_int * 3),
        ("min_vertex_distance", ctypes.c_double),
        ("h", ctypes.c_double * 3),
    ]


class DgSolverControl(ctypes.Structure):
    _fields_ = [
        ("max_iterations", ctypes.c_int),
        ("abs_tolerance", ctypes.c_double),
        ("rel_tolerance", ctypes.c_double),
        ("restart", ctypes.c_int),
        ("flexible", ctypes.c_int),
        ("prec_precision", ctypes.c_int),
        ("inner_iterations", ctypes.c_int),
        ("cell_solver", ctypes.c_int),
        ("iterations", ctypes.c_int),
        ("initial_residual", ctypes.c_double),
        ("final_residual", ctypes.c_double),
    ]


class DgServeRequest(ctypes.Structure):
    _fields_ = [
        ("degree", ctypes.c_int),
        ("cells_per_direction", ctypes.c_int),
        ("velocity", ctypes.c_double * 3),
        ("center", ctypes.c_double * 3),
        ("width", ctypes.c_double),
        ("final_time", ctypes.c_double),
        ("courant", ctypes.c_double),
    ]


(DG_REQUEST_QUEUED, DG_REQUEST_ACTIVE, DG_REQUEST_DONE, DG_REQUEST_UNKNOWN,
 DG_REQUEST_FAILED) = range(5)

_vp = ctypes.c_void_p
_d = ctypes.c_double
_i = ctypes.c_int
_i64 = ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); every symbol of include/dgops.h
SIGNATURES = {
    "dg_last_error": (ctypes.c_char_p, []),
    "dg_config_default": (None, [ctypes.POINTER(DgConfig)]),
    "dg_create": (_i, [ctypes.POINTER(DgConfig), _vp, ctypes.POINTER(_vp)]),
    "dg_destroy": (_i, [_vp]),
    "dg_get_info": (_i, [_vp, ctypes.POINTER(DgInfo)]),
    "dg_set_stream": (_i, [_vp, _vp]),
    "dg_apply": (_i, [_vp, _vp, _vp, _d]),
    "dg_perform_stage": (_i, [_vp, _d, _d, _d, _vp, _vp, _vp, _vp]),
    "dg_lsrk_time_step": (_i, [_vp, _i, _d, _d, _vp, _vp, _vp]),
    "dg_lsrk_time_steps": (_i, [_vp, _i, _d, _d, _i, _vp, _vp, _vp]),
    "dg_set_step_launch": (_i, [_vp, _i]),
    "dg_get_step_launch": (_i, [_vp, ctypes.POINTER(_i), _dp, _dp]),
    "dg_lsrk_n_stages": (_i, [_i]),
    "dg_lsrk_coefficients": (_i, [_i, _dp, _dp]),
    "dg_project_initial": (_i, [_vp, _vp, _d]),
    "dg_compute_mass_and_energy": (_i, [_vp, _vp, _dp]),
    "dg_l2_error": (_i, [_vp, _vp, _d, _dp]),
    "dg_write_vtu": (_i, [_vp, _vp, _d, ctypes.c_char_p]),
    "dg_host_write_vtu": (_i, [ctypes.POINTER(DgConfig), _dp, _d,
                               ctypes.c_char_p]),
    "dg_set_time": (_i, [_vp, _d, _d]),
    "dg_vmult": (_i, [_vp, _vp, _vp]),
    "dg_apply_inverse_mass_matrix": (_i, [_vp, _vp]),
    "dg_precondition_block_jacobi": (_i, [_vp, _vp, _vp]),
    "dg_irk_compute_rhs": (_i, [_vp, _vp, _vp]),
    "dg_irk_vmult": (_i, [_vp, _vp, _vp]),
    "dg_irk_update_solution": (_i, [_vp, _vp, _vp]),
    "dg_irk_precondition_block_jacobi": (_i, [_vp, _vp, _vp]),
    "dg_irk_precondition_cell_solver": (_i, [_vp, _vp, _vp]),
    "dg_irk_cell_operator": (_i, [_vp, _vp, _vp]),
    "dg_irk_coefficients": (_i, [_i, _dp, _dp, _dp]),
    "dg_irk_solve": (_i, [_vp, _vp, _vp, ctypes.POINTER(DgSolverControl)]),
    "dg_irk_block_jacobi_vmult": (_i, [_vp, ctypes.POINTER(DgSolverControl),
                                       _vp, _vp]),
    "dg_irk_time_step": (_i, [_vp, _d, _d, _vp,
                              ctypes.POINTER(DgSolverControl)]),
    "dg_dirk_time_step": (_i, [_vp, _i, _d, _d, _vp,
                               ctypes.POINTER(DgSolverControl)]),
    "dg_irk_integrator_create": (_i, [_vp, ctypes.POINTER(_vp)]),
    "dg_irk_integrator_destroy": (_i, [_vp]),
    "dg_irk_integrator_step": (_i, [_vp, _d, _d, _vp,
                                    ctypes.POINTER(DgSolverControl), _dp]),
    "dg_set_tile_shape": (_i, [_vp, _i]),
    "dg_get_tile_shape": (_i, [_vp, ctypes.POINTER(_i), _dp, _dp]),
    "dg_vector_dot": (_i, [_vp, _i, _i64, _vp, _vp, _dp]),
    "dg_vector_multi_dot": (_i, [_vp, _i, _i64, _i, _vp, _vp, _dp]),
    "dg_vector_multi_dot_noalloc": (_i, [_vp, _i, _i64, _i, _vp, _vp,
                                         _dp]),
    "dg_vector_multi_axpy": (_i, [_vp, _i, _i64, _i, _dp, _vp, _vp]),
    "dg_vector_axpby": (_i, [_vp, _i, _i64, _d, _vp, _d, _vp]),
    "dg_vector_copy_convert": (_i, [_vp, _i64, _i, _vp, _i, _vp]),
    "dg_export_cell_coordinates": (_i, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "dg_export_neighbors": (_i, [_vp, ctypes.POINTER(ctypes.c_int32)]),
    "dg_export_cell_vertices": (_i, [_vp, _dp]),
    "dg_serve_request_config": (_i, [ctypes.POINTER(DgServeRequest),
                                     ctypes.POINTER(DgConfig), _dp,
                                     ctypes.POINTER(_i)]),
    "dg_server_create": (_i, [_i64, _vp, ctypes.POINTER(_vp)]),
    "dg_server_destroy": (_i, [_vp]),
    "dg_server_submit": (_i, [_vp, ctypes.POINTER(DgServeRequest),
                              ctypes.POINTER(_i64)]),
    "dg_server_step": (_i, [_vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i),
                            ctypes.POINTER(_i)]),
    "dg_server_request_info": (_i, [_vp, _i64, ctypes.POINTER(_i),
                                    ctypes.POINTER(_i64), ctypes.POINTER(_i),
                                    ctypes.POINTER(_i)]),
    "dg_server_take_result": (_i, [_vp, _i64, _vp]),
    "dg_server_memory": (_i, [_vp, ctypes.POINTER(_i64)]),
    "dg_host_index_map": (_i, [ctypes.POINTER(DgConfig),
                               ctypes.POINTER(ctypes.c_int64), _dp,
                               ctypes.POINTER(ctypes.c_int64)]),
}

_lib = None


class DgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dgops error {code}: {msg}")
        self.code = code


def load(path=LIB_PATH):
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with "
            "`python -m advection_miniapp_amd.build` (there is no CPU "
            "fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(code):
    if code != DG_OK:
        msg = load().dg_last_error()
        raise DgError(code, msg.decode() if msg else "")
