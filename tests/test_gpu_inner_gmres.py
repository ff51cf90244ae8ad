"""The inner GMRES of the mixed-precision IRK solve (device-resident
scalars, fixed step count) against the outer FGMRES on the host.

m steps of right-preconditioned GMRES from zero give the minimizer x_m of
|b - A x| over the preconditioned Krylov space. An outer FGMRES limited to
one iteration whose preconditioner is that inner solve returns alpha x_m
with alpha minimizing |b - alpha A x_m|, and alpha = 1 because x_m is
optimal along its own direction. So one outer step around an inner
GMRES(m) equals m steps of the outer FGMRES with the same preconditioner
(advection_solver_irk_real.cc:2725-2757 for the solver setup)."""
import pytest

from advection_miniapp_amd import AdvectionOperation
from advection_miniapp_amd import _lib as L
from advection_miniapp_amd.operators import solver_control
from helpers import Cfg

pytestmark = pytest.mark.gpu


def stage_rhs(c, dt):
    op = AdvectionOperation(c.dg())
    u = op.initialize_dof_vector()
    op.project_initial(u, 0.0)
    op.set_time(0.0, dt)
    rhs = op.initialize_dof_vector(blocks=c.irk_stages)
    op.compute_rhs(rhs, u)
    return op, rhs


def capped_solve(op, rhs, **kw):
    """The outer solver's iterate after exactly max_iterations steps (it
    reports no convergence at that cap)."""
    z = op.initialize_dof_vector(blocks=op.n_stages)
    ctl = solver_control(rel_tolerance=1e-30, **kw)
    try:
        op.irk_solve(z, rhs, ctl)
    except L.DgError as e:
        assert e.code == L.DG_ERR_NO_CONVERGENCE, e
    assert ctl.iterations == kw["max_iterations"]
    return z


@pytest.mark.parametrize("m", [1, 3, 4, 8])
@pytest.mark.parametrize("case", ["3d_k4_pert", "2d_k3_vortex"])
def test_inner_gmres_equals_outer_fgmres_steps(gpu_lib, case, m):
    if case == "3d_k4_pert":
        c = Cfg(dim=3, degree=4, refine=2, perturbation=0.1, velocity_kind=1,
                factor_skew=0.5, irk_stages=3, solution_kind=3,
                solution_param=[0.5, 0.5, 0.5, 0.1])
        dt = 0.2
    else:
        c = Cfg(dim=2, degree=3, refine=3, velocity_kind=3, time_factor=1,
                final_time=8.0, factor_skew=0.5, irk_stages=2,
                solution_kind=5)
        dt = 0.05
    op, rhs = stage_rhs(c, dt)
    one = capped_solve(op, rhs, max_iterations=1, restart=30,
                       prec_precision=L.DG_FP64, inner_iterations=m)
    ref = capped_solve(op, rhs, max_iterations=m, restart=m,
                       prec_precision=L.DG_FP64)
    err = ((one - ref).norm() / ref.norm()).item()
    assert err < 1e-9, err
    # the FP32 inner solve (FP32 vmult and FDM): the same iterate up to
    # the FP32 operators' error, which the FDM's eigenvector transforms
    # (condition numbers up to 1e4) amplify; measured 2e-4 .. 1.5e-3
    one32 = capped_solve(op, rhs, max_iterations=1, restart=30,
                         prec_precision=L.DG_FP32, inner_iterations=m)
    err32 = ((one32 - ref).norm() / ref.norm()).item()
    assert err32 < 5e-3, err32


def test_inner_gmres_rejects_more_than_16_steps(gpu_lib):
    """Checked before any work, whatever the right-hand side."""
    c = Cfg(dim=2, degree=2, refine=2, velocity=[1.0, 0.3], factor_skew=0.5,
            irk_stages=2)
    op, rhs = stage_rhs(c, 0.1)
    z = op.initialize_dof_vector(blocks=2)
    with pytest.raises(L.DgError) as e:
        op.irk_solve(z, rhs, solver_control(inner_iterations=17,
                                            prec_precision=L.DG_FP32))
    assert e.value.code == L.DG_ERR_UNSUPPORTED
