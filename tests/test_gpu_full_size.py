"""Properties at the full sizes of the benchmark configurations (too large
for the oracle): configuration 2 (3D, degree 4, 64^3 perturbed cells,
cellular speed, 32,768,000 DoF) and configuration 4 (3D, degree 4, 32^3
cells, Radau IIA(3), 12,288,000 unknowns) with its Burgers add-on (degree 3,
48^3 cells, 21,233,664 unknowns), and configuration 3 (SIPG, 32^3 cells,
degrees 3, 5, 7)."""
import numpy as np
import pytest
import torch

from advection_miniapp_amd import (AdvectionOperation,
                                   LowStorageRungeKuttaIntegrator)
from advection_miniapp_amd import _lib as L
from advection_miniapp_amd.operators import solver_control
from helpers import Cfg

pytestmark = pytest.mark.gpu


def config2(**kw):
    base = dict(dim=3, degree=4, refine=6, perturbation=0.1, seed=42,
                velocity_kind=1, solution_kind=3,
                solution_param=[0.5, 0.5, 0.5, 0.1], lo=[0, 0, 0],
                hi=[1, 1, 1])
    base.update(kw)
    return Cfg(**base)


@pytest.fixture(scope="module")
def op2(gpu_lib):
    op = AdvectionOperation(config2().dg())
    assert op.n_dofs == 32768000
    yield op
    del op
    torch.cuda.empty_cache()


def test_config2_apply_conserves_mass_and_is_linear(op2):
    """Periodic, conservative form: the integral of M^-1(-L u) vanishes
    (the numerical fluxes cancel between neighbours and the cell term
    integrates grad 1 = 0); the operator is linear."""
    u = op2.initialize_dof_vector()
    op2.project_initial(u, 0.0)
    g = torch.Generator(device="cuda").manual_seed(3)
    v = torch.randn(u.numel(), dtype=u.dtype, device="cuda", generator=g)
    du, dv = op2.initialize_dof_vector(), op2.initialize_dof_vector()
    op2.apply(u, du, 0.0)
    op2.apply(v, dv, 0.0)
    for d in (du, dv):
        mass, energy, _ = op2.compute_mass_and_energy(d)
        assert abs(mass) < 1e-11 * energy ** 0.5, (mass, energy)
    w = 2.0 * u - 0.75 * v
    dw = op2.initialize_dof_vector()
    op2.apply(w, dw, 0.0)
    ref = 2.0 * du - 0.75 * dv
    assert ((dw - ref).abs().max() / ref.abs().max()).item() < 1e-12


def test_config2_fused_step_equals_two_pass_stages(op2):
    """Three distinct registers run each LSRK(5,4) stage as one fused
    kernel; the reference's aliasing pattern (perform_stage with
    current_ri == solution, next_ri == current_ri) takes the two-pass
    path. Both give the same step."""
    import ctypes
    u = op2.initialize_dof_vector()
    op2.project_initial(u, 0.0)
    a = u.clone()
    r1, r2 = op2.initialize_dof_vector(), op2.initialize_dof_vector()
    dt = 2e-4
    LowStorageRungeKuttaIntegrator().perform_time_step(op2, 0.0, dt, a, r1,
                                                       r2)
    # the reference's integrator loop (advection_solver.cc:906-939)
    bi, ai = (ctypes.c_double * 16)(), (ctypes.c_double * 16)()
    ns = L.load().dg_lsrk_coefficients(L.DG_LSRK_5_4, bi, ai)
    assert ns == 5
    b = u.clone()
    ri, ki = op2.initialize_dof_vector(), op2.initialize_dof_vector()
    op2.perform_stage(0.0, bi[0] * dt, ai[0] * dt, b, ri, b, ri)
    sum_previous_bi = 0.0
    for stage in range(1, ns):
        c_i = sum_previous_bi + ai[stage - 1]
        op2.perform_stage(c_i * dt, bi[stage] * dt,
                          0.0 if stage == ns - 1 else ai[stage] * dt,
                          ri, ki, b, ri)
        sum_previous_bi += bi[stage - 1]
    assert ((a - b).abs().max() / b.abs().max()).item() < 1e-13


def test_config4_irk_solve_reaches_tolerance_at_full_size(gpu_lib):
    c = Cfg(dim=3, degree=4, refine=5, velocity_kind=1, factor_skew=0.5,
            irk_stages=3, solution_kind=3,
            solution_param=[0.5, 0.5, 0.5, 0.1])
    op = AdvectionOperation(c.dg())
    s = 3
    assert op.n_dofs * s == 12288000
    u = op.initialize_dof_vector()
    op.project_initial(u, 0.0)
    dt = 5.0 / 32 / 2.0
    op.set_time(0.0, dt)
    rhs = op.initialize_dof_vector(blocks=s)
    op.compute_rhs(rhs, u)
    for kw in (dict(prec_precision=L.DG_FP64),
               dict(prec_precision=L.DG_FP32, inner_iterations=4)):
        z = op.initialize_dof_vector(blocks=s)
        ctl = solver_control(max_iterations=400, rel_tolerance=1e-10,
                             restart=30, **kw)
        op.irk_solve(z, rhs, ctl)
        res = op.initialize_dof_vector(blocks=s)
        op.vmult_irk(res, z)
        res -= rhs
        assert (res.norm() / rhs.norm()).item() < 1e-10, kw


def burgers_exact(x, t):
    """u = -sin(pi (x - 2 u t)): the first component of the Burgers flux
    u (x) u moves at speed 2u; Newton per point below the break time
    1/(2 pi)."""
    u = -torch.sin(np.pi * x)
    for _ in range(60):
        a = np.pi * (x - 2.0 * u * t)
        u = u - (u + torch.sin(a)) / (1.0 - 2.0 * np.pi * t * torch.cos(a))
    return u


def test_config4_burgers_converges_to_characteristic_solution(gpu_lib):
    """The Burgers add-on at its own size (degree 3, 48^3 cells, 21.2 M
    unknowns) and at 24^3: LSRK(5,4) runs from u0 = (-sin pi x, 0, 0) to
    t = 0.05 against the exact solution by characteristics; the error
    falls at order > 3 from 24^3 to 48^3 and the other two components stay
    zero."""
    from helpers import nodal_points
    T, k, errs = 0.05, 3, []
    for refine in (3, 4):
        c = Cfg(dim=3, degree=k, n_components=3, operator_kind=2,
                n_coarse=[3, 3, 3], refine=refine, lo=[-1, -1, -1],
                hi=[1, 1, 1], solution_kind=6)
        op = AdvectionOperation(c.dg())
        print(f"burgers {3 * 2 ** refine}^3: {op.vector_size} unknowns",
              flush=True)
        u = op.initialize_dof_vector()
        op.project_initial(u, 0.0)
        r1, r2 = op.initialize_dof_vector(), op.initialize_dof_vector()
        h = 2.0 / (3 * 2 ** refine)
        n = int(np.ceil(T / (0.25 * h / (k + 1) ** 1.5)))
        integ = LowStorageRungeKuttaIntegrator(L.DG_LSRK_5_4)
        integ.perform_time_steps(op, 0.0, T / n, n, u, r1, r2)
        uc = u.view(-1, 3, (k + 1) ** 3)
        x = torch.from_numpy(np.ascontiguousarray(
            nodal_points(op.cell_vertices(), k + 1, 3)[..., 0])).cuda()
        errs.append((uc[:, 0, :] - burgers_exact(x, T)).abs().max().item())
        assert uc[:, 1:, :].abs().max().item() == 0.0
        del op, u, r1, r2, uc, x
        torch.cuda.empty_cache()
    rate = np.log2(errs[0] / errs[1])
    # measured: 4.0e-5 and 3.3e-6, rate 3.59
    print(f"max nodal errors {errs}, rate {rate:.2f}", flush=True)
    assert errs[1] < 1e-4 and rate > 3.0, (errs, rate)


def test_config3_sipg_runs_match_the_decaying_exact_solution(gpu_lib):
    """Configuration 3 at its own size (32^3 cells, degrees 3, 5, 7, up to
    16.8 M DoF): LSRK(5,4) runs of advection + SIPG from the product of
    sines to t = 0.01 against the exact solution, which is the same
    product translated by a t and damped by exp(-3 nu (2 pi)^2 t); the
    relative L2 error falls with the degree."""
    import math
    T, errs = 0.01, []
    for k in (3, 5, 7):
        c = Cfg(dim=3, degree=k, refine=5, velocity=[1.0, 0.5, 0.25],
                operator_kind=1, viscosity=1e-3,
                penalty_factor=(k + 1) ** 2, solution_kind=2,
                solution_param=[2 * math.pi, 1e-3])
        op = AdvectionOperation(c.dg())
        print(f"sipg k={k}: {op.n_dofs} DoF", flush=True)
        u = op.initialize_dof_vector()
        op.project_initial(u, 0.0)
        r1, r2 = op.initialize_dof_vector(), op.initialize_dof_vector()
        n = int(math.ceil(T / (0.1 / 32 / (k + 1) ** 2)))
        LowStorageRungeKuttaIntegrator(L.DG_LSRK_5_4).perform_time_steps(
            op, 0.0, T / n, n, u, r1, r2)
        err, nrm = op.l2_error(u, T)
        errs.append(err / nrm)
        del op, u, r1, r2
        torch.cuda.empty_cache()
    # measured: 7.7e-7, 7.4e-11, 3.3e-15
    print(f"relative L2 errors {errs}", flush=True)
    assert errs[0] > errs[1] > errs[2], errs
    assert errs[0] < 5e-6 and errs[1] < 1e-9 and errs[2] < 1e-12, errs
