"""Python mirror of the reference's operator and integrator interfaces.

The classes keep the names, argument order and meaning of the C++ classes
of kronbichler/advection_miniapp so that driver code reads the same; all
work runs in libdgops.so (HIP, gfx950) on device tensors. There is no CPU
fallback: a missing library or a CPU tensor raises.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L

_DTYPES = {L.DG_FP64: torch.float64, L.DG_FP32: torch.float32}


def make_config(**kw):
    """DgConfig from dg_config_default() with the given fields replaced.

    Array fields accept sequences; missing trailing entries keep the
    default."""
    lib = L.load()
    cfg = L.DgConfig()
    lib.dg_config_default(ctypes.byref(cfg))
    for k, v in kw.items():
        if not hasattr(cfg, k):
            raise KeyError(f"unknown config field {k}")
        cur = getattr(cfg, k)
        if isinstance(cur, ctypes.Array):
            for i, x in enumerate(v):
                cur[i] = x
        else:
            setattr(cfg, k, v)
    return cfg


def _ptr(t, dtype, numel, name):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name}: vector must live on the GPU")
    if t.dtype != dtype:
        raise TypeError(f"{name}: dtype {t.dtype}, expected {dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: vector must be contiguous")
    if t.numel() < numel:
        raise ValueError(f"{name}: {t.numel()} entries, need {numel}")
    return ctypes.c_void_p(t.data_ptr())


class LowStorageRungeKuttaIntegrator:
    """advection_solver.cc:823-944: two-register low-storage Runge-Kutta."""

    def __init__(self, scheme=L.DG_LSRK_5_4):
        self.scheme = scheme
        self._stages = L.load().dg_lsrk_n_stages(scheme)
        if self._stages < 0:
            raise ValueError("unknown LSRK scheme")

    def n_stages(self):
        return self._stages

    def perform_time_step(self, pde_operator, current_time, time_step,
                          solution, vec_ri, vec_ki):
        pde_operator._lsrk_time_step(self.scheme, current_time, time_step,
                                     solution, vec_ri, vec_ki)

    def perform_time_steps(self, pde_operator, current_time, time_step,
                           n_steps, solution, vec_ri, vec_ki):
        """n_steps calls of perform_time_step (the driver's time loop,
        advection_solver.cc:1264-1299) in one library call: for an operator
        without time dependence the steps replay one captured HIP graph."""
        pde_operator._lsrk_time_steps(self.scheme, current_time, time_step,
                                      n_steps, solution, vec_ri, vec_ki)


class AdvectionOperation:
    """The reference's AdvectionOperation (explicit: advection_solver.cc:330-421;
    implicit: advection_solver_implicit_real.cc:816-909; IRK:
    advection_solver_irk_real.cc:1227-1331) on one GPU."""

    def __init__(self, config=None, stream=None, **kw):
        self._lib = L.load()
        if config is None:
            config = make_config(**kw)
        elif kw:
            raise ValueError("give either a DgConfig or keyword fields")
        self.config = config
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self._stream = stream
        h = ctypes.c_void_p()
        L.check(self._lib.dg_create(ctypes.byref(config),
                                    ctypes.c_void_p(stream), ctypes.byref(h)))
        self._h = h
        info = L.DgInfo()
        L.check(self._lib.dg_get_info(h, ctypes.byref(info)))
        self.info = info
        self.n_cells = info.n_cells
        self.n_dofs = info.n_dofs
        self.vector_size = info.vector_size
        self.dofs_per_cell = info.dofs_per_cell
        self.min_vertex_distance = info.min_vertex_distance
        self.dtype = _DTYPES[config.precision]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self._lib.dg_destroy(h)
            except Exception:
                pass
            self._h = None

    # ---- vectors --------------------------------------------------------
    def initialize_dof_vector(self, blocks=1, dtype=None):
        return torch.zeros(self.vector_size * blocks,
                           dtype=dtype or self.dtype, device="cuda")

    def _v(self, t, name, blocks=1, dtype=None):
        return _ptr(t, dtype or self.dtype, self.vector_size * blocks, name)

    # ---- explicit operator ------------------------------------------------
    def apply(self, src, dst, current_time):
        L.check(self._lib.dg_apply(self._h, self._v(src, "src"),
                                   self._v(dst, "dst"), current_time))

    def perform_stage(self, current_time, factor_solution, factor_ai,
                      current_ri, vec_ki, solution, next_ri):
        L.check(self._lib.dg_perform_stage(
            self._h, current_time, factor_solution, factor_ai,
            self._v(current_ri, "current_ri"), self._v(vec_ki, "vec_ki"),
            self._v(solution, "solution"), self._v(next_ri, "next_ri")))

    def _lsrk_time_step(self, scheme, t, dt, solution, vec_ri, vec_ki):
        L.check(self._lib.dg_lsrk_time_step(
            self._h, scheme, t, dt, self._v(solution, "solution"),
            self._v(vec_ri, "vec_ri"), self._v(vec_ki, "vec_ki")))

    def _lsrk_time_steps(self, scheme, t, dt, n_steps, solution, vec_ri,
                         vec_ki):
        L.check(self._lib.dg_lsrk_time_steps(
            self._h, scheme, t, dt, n_steps, self._v(solution, "solution"),
            self._v(vec_ri, "vec_ri"), self._v(vec_ki, "vec_ki")))

    def set_step_launch(self, mode=0):
        """Launch form of LowStorageRungeKuttaIntegrator.perform_time_steps:
        0 measured on the first long call, 1 stream launches, 2 graph
        replay wherever the operator allows it."""
        L.check(self._lib.dg_set_step_launch(self._h, mode))

    def step_launch(self):
        """The launch form in use (0 before it was measured) and the
        measured time of DG_GRAPH_STEPS steps in each form (ms)."""
        m = ctypes.c_int()
        e, g = ctypes.c_double(), ctypes.c_double()
        L.check(self._lib.dg_get_step_launch(self._h, ctypes.byref(m),
                                             ctypes.byref(e),
                                             ctypes.byref(g)))
        return dict(mode=m.value, eager_ms=e.value, graph_ms=g.value)

    def set_tile_shape(self, cells_per_workgroup=0):
        """Force the operator kernels' tile shape; 0 = autotuned."""
        L.check(self._lib.dg_set_tile_shape(self._h, cells_per_workgroup))

    def tile_shape(self):
        """The tile shape of the last operator launch: cells per
        workgroup, the tuned time (ms, 0 if not tuned) and the tuner's
        time of every shape {cells per workgroup: ms, -1 if not timed}."""
        c = ctypes.c_int()
        ms = ctypes.c_double()
        cand = (ctypes.c_double * 6)()
        L.check(self._lib.dg_get_tile_shape(self._h, ctypes.byref(c),
                                            ctypes.byref(ms), cand))
        return dict(cells_per_workgroup=c.value, tuned_ms=ms.value,
                    candidates_ms={int(cand[2 * i]): cand[2 * i + 1]
                                   for i in range(3)})

    def project_initial(self, dst, time=0.0):
        L.check(self._lib.dg_project_initial(self._h, self._v(dst, "dst"),
                                             time))

    def compute_mass_and_energy(self, vec):
        out = (ctypes.c_double * 3)()
        L.check(self._lib.dg_compute_mass_and_energy(
            self._h, self._v(vec, "vec"), out))
        return tuple(out)

    def l2_error(self, vec, time):
        """(||u_h - u(t)||, ||u(t)||) as output_results computes them."""
        out = (ctypes.c_double * 2)()
        L.check(self._lib.dg_l2_error(self._h, self._v(vec, "vec"), time,
                                      out))
        return tuple(out)

    def write_vtu(self, vec, time, filename):
        """VTU file with "solution" and "analytic_solution" at `time`, one
        Lagrange cell of the operator's degree per cell, as output_results
        writes it (advection_solver.cc:1217-1241)."""
        L.check(self._lib.dg_write_vtu(self._h, self._v(vec, "vec"), time,
                                       str(filename).encode()))

    # ---- implicit operator -------------------------------------------------
    def set_time(self, current_time, time_step):
        L.check(self._lib.dg_set_time(self._h, current_time, time_step))

    def vmult(self, dst, src):
        L.check(self._lib.dg_vmult(self._h, self._v(dst, "dst"),
                                   self._v(src, "src")))

    def apply_inverse_mass_matrix(self, dst):
        L.check(self._lib.dg_apply_inverse_mass_matrix(
            self._h, self._v(dst, "dst")))

    def precondition_block_jacobi(self, dst, src):
        L.check(self._lib.dg_precondition_block_jacobi(
            self._h, self._v(dst, "dst"), self._v(src, "src")))

    # ---- IRK -----------------------------------------------------------------
    @property
    def n_stages(self):
        return self.config.irk_stages

    def compute_rhs(self, dst, src):
        s = self.n_stages
        L.check(self._lib.dg_irk_compute_rhs(
            self._h, self._v(dst, "dst", s), self._v(src, "src")))

    def vmult_irk(self, dst, src):
        s = self.n_stages
        L.check(self._lib.dg_irk_vmult(self._h, self._v(dst, "dst", s),
                                       self._v(src, "src", s)))

    def update_solution(self, solution, stage_solution):
        s = self.n_stages
        L.check(self._lib.dg_irk_update_solution(
            self._h, self._v(solution, "solution"),
            self._v(stage_solution, "stage_solution", s)))

    def precondition_block_jacobi_irk(self, dst, src):
        s = self.n_stages
        L.check(self._lib.dg_irk_precondition_block_jacobi(
            self._h, self._v(dst, "dst", s), self._v(src, "src", s)))

    def precondition_cell_solver_irk(self, dst, src):
        """Block-Jacobi step with the batched cell solver (one GMRES step
        per cell in the collocation basis, advection_solver_irk_real.cc:
        2108-2126)."""
        s = self.n_stages
        L.check(self._lib.dg_irk_precondition_cell_solver(
            self._h, self._v(dst, "dst", s), self._v(src, "src", s)))

    def cell_operator_irk(self, dst, src):
        """The cell-local stage operator at the Gauss points (values at the
        Gauss points in and out, neighbour traces zero)."""
        s = self.n_stages
        L.check(self._lib.dg_irk_cell_operator(
            self._h, self._v(dst, "dst", s), self._v(src, "src", s)))

    def block_jacobi_vmult(self, dst, src, control):
        """The preconditioner irk_solve applies for `control` (FDM in the
        control's precision on FP64 block vectors, cell solver or inner
        GMRES): BlockJacobi::vmult, advection_solver_irk_real.cc:2462-2479."""
        s = self.n_stages
        L.check(self._lib.dg_irk_block_jacobi_vmult(
            self._h, ctypes.byref(control), self._v(dst, "dst", s),
            self._v(src, "src", s)))

    def irk_solve(self, stage, rhs, control):
        """(F)GMRES on the stage system from the initial guess in `stage`
        (SolverFGMRES::solve, advection_solver_irk_real.cc:2729-2757);
        `control` (dg_solver_control) returns iterations and residuals."""
        s = self.n_stages
        L.check(self._lib.dg_irk_solve(self._h, self._v(stage, "stage", s),
                                       self._v(rhs, "rhs", s),
                                       ctypes.byref(control)))
        return control

    # ---- index data -----------------------------------------------------------
    def cell_coordinates(self):
        a = np.zeros((self.n_cells, 3), dtype=np.int32)
        L.check(self._lib.dg_export_cell_coordinates(
            self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return a

    def neighbors(self):
        a = np.zeros((self.n_cells, 2 * self.config.dim), dtype=np.int32)
        L.check(self._lib.dg_export_neighbors(
            self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return a

    def cell_vertices(self):
        d = self.config.dim
        a = np.zeros((self.n_cells, 2 ** d, d), dtype=np.float64)
        L.check(self._lib.dg_export_cell_vertices(
            self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return a


def solver_control(max_iterations=400, rel_tolerance=1e-8, abs_tolerance=0.0,
                   restart=10, flexible=True, prec_precision=L.DG_FP64,
                   inner_iterations=0, cell_solver=False):
    """dg_solver_control with the reference's defaults
    (SolverControl(400, 1e-8 |rhs|), max_basis_size 10, FGMRES)."""
    c = L.DgSolverControl()
    c.max_iterations = max_iterations
    c.rel_tolerance = rel_tolerance
    c.abs_tolerance = abs_tolerance
    c.restart = restart
    c.flexible = 1 if flexible else 0
    c.prec_precision = prec_precision
    c.inner_iterations = inner_iterations
    c.cell_solver = 1 if cell_solver else 0
    return c


class ImplicitRungeKuttaIntegrator:
    """Radau IIA stepping of advection_solver_irk_real.cc:2484-2779 with the
    extrapolated initial guess of the last four stage solutions."""

    def __init__(self, operator):
        self.op = operator
        h = ctypes.c_void_p()
        L.check(operator._lib.dg_irk_integrator_create(operator._h,
                                                       ctypes.byref(h)))
        self._h = h
        self.project_sol = np.zeros(4)

    def __del__(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.op._lib.dg_irk_integrator_destroy(self._h)
            self._h = None

    def perform_time_step(self, current_time, time_step, solution,
                          control=None):
        control = control or solver_control()
        p = (ctypes.c_double * 4)()
        L.check(self.op._lib.dg_irk_integrator_step(
            self._h, current_time, time_step, self.op._v(solution, "solution"),
            ctypes.byref(control), p))
        self.project_sol = np.array(p)
        return control


class DiagonallyImplicitRungeKuttaIntegrator:
    """advection_solver_implicit_real.cc:1561-1748 (1, 3, 4 or 7 stages)."""

    def __init__(self, n_stages, operator):
        self.n_stages = n_stages
        self.op = operator

    def perform_time_step(self, solution, current_time, time_step,
                          control=None):
        control = control or solver_control(max_iterations=1000)
        L.check(self.op._lib.dg_dirk_time_step(
            self.op._h, self.n_stages, current_time, time_step,
            self.op._v(solution, "solution"), ctypes.byref(control)))
        return control


def irk_coefficients(stages):
    A = (ctypes.c_double * (stages * stages))()
    b = (ctypes.c_double * stages)()
    c = (ctypes.c_double * stages)()
    L.check(L.load().dg_irk_coefficients(stages, A, b, c))
    return (np.array(A).reshape(stages, stages), np.array(b), np.array(c))


def explicit_time_step(min_vertex_distance, speed_norm, degree, courant,
                       final_time):
    """dt = C h_min / |a| / max(k, 1)^1.5, shortened so that final_time is a
    whole number of steps (advection_solver.cc:1163-1167)."""
    dt = courant * min_vertex_distance / speed_norm / \
        max(1, degree) ** 1.5
    return final_time / np.ceil(final_time / dt)


def implicit_time_step(min_vertex_distance, courant, final_time):
    """dt = C h_min / 2, shortened so that final_time is a whole number of
    steps (advection_solver_irk_real.cc:2352-2354)."""
    dt = courant * min_vertex_distance / 2.0
    return final_time / np.ceil(final_time / dt)
