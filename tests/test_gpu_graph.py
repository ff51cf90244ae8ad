"""dg_lsrk_time_steps: several LSRK steps in one call. For an operator that
does not depend on time, runs of DG_GRAPH_STEPS steps may replay a captured
HIP graph; every result must equal the same steps taken one by one with
dg_lsrk_time_step, bit for bit (same kernels, same arguments), in each
launch form, also when graph replays and eager steps alternate on the same
registers."""
import numpy as np
import pytest
import torch

from advection_miniapp_amd import AdvectionOperation, \
    LowStorageRungeKuttaIntegrator
from advection_miniapp_amd import _lib as L
from helpers import Cfg

pytestmark = pytest.mark.gpu

G = L.DG_GRAPH_STEPS

CASES = {
    # periodic, constant speed: graph replay allowed
    "2d_k3_config1": dict(dim=2, degree=3, refine=5, lo=[-1, -1],
                          hi=[1, 1], velocity=[1.0, 0.15], solution_kind=2,
                          solution_param=[np.pi, 0.0]),
    "3d_k4_pert_cellular": dict(dim=3, degree=4, refine=3, perturbation=0.1,
                                velocity_kind=1, solution_kind=3,
                                solution_param=[0.5, 0.5, 0.5, 0.1]),
    "3d_burgers_k3": dict(dim=3, degree=3, n_components=3, operator_kind=2,
                          refine=2, lo=[-1, -1, -1], hi=[1, 1, 1],
                          solution_kind=6),
    # SIPG: the fused kernel on axis-aligned cells; on perturbed cells
    # three launches, one of them reading the face table built on first use
    "3d_k3_sipg": dict(dim=3, degree=3, refine=2, velocity=[1.0, 0.5, 0.25],
                       operator_kind=1, viscosity=1e-2, penalty_factor=16,
                       solution_kind=2, solution_param=[2 * np.pi, 1e-2]),
    "3d_k3_sipg_pert": dict(dim=3, degree=3, refine=2, perturbation=0.1,
                            seed=7, velocity=[1.0, 0.5, 0.25],
                            operator_kind=1, viscosity=5e-2,
                            penalty_factor=16, solution_kind=2,
                            solution_param=[2 * np.pi, 5e-2]),
    # Gauss-Lobatto switches: the same kernels with GL tables, and for the
    # mass matrix alone the two-pass stage (integrate, inverse mass, update)
    "3d_k3_gl_quad": dict(dim=3, degree=3, refine=2, perturbation=0.1,
                          velocity_kind=1, solution_kind=3,
                          solution_param=[0.5, 0.5, 0.5, 0.1], use_gl_quad=1),
    "3d_k3_gl_quad_mass": dict(dim=3, degree=3, refine=2, perturbation=0.1,
                               velocity_kind=1, solution_kind=3,
                               solution_param=[0.5, 0.5, 0.5, 0.1],
                               use_gl_quad_mass=1),
    # time-dependent speed, and inflow boundaries: steps launched one by one
    "2d_k4_vortex_time": dict(dim=2, degree=4, refine=3, perturbation=0.1,
                              velocity_kind=2, time_factor=1,
                              final_time=8.0, solution_kind=5),
    "3d_k2_dirichlet": dict(dim=3, degree=2, n_coarse=[2, 1, 1], refine=1,
                            periodic=[0, 1, 0], velocity=[1.1, 0.15, -0.05],
                            solution_kind=1),
}
AUTONOMOUS = {"2d_k3_config1", "3d_k4_pert_cellular", "3d_burgers_k3",
              "3d_k3_sipg", "3d_k3_sipg_pert", "3d_k3_gl_quad",
              "3d_k3_gl_quad_mass"}


def run_eager(op, integ, u, r1, r2, t, dt, n):
    for s in range(n):
        integ.perform_time_step(op, t + s * dt, dt, u, r1, r2)


def vectors(op, u0):
    return u0.clone(), op.initialize_dof_vector(), op.initialize_dof_vector()


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("mode", [0, 1, 2])
def test_time_steps_equal_single_steps(gpu_lib, name, mode):
    c = Cfg(**CASES[name])
    op = AdvectionOperation(c.dg())
    op.set_step_launch(mode)
    print(f"case {name} launch form {mode}", flush=True)
    integ = LowStorageRungeKuttaIntegrator(1)
    dt = 1e-3
    n1, n2 = 3 * G + 3, 2 * G + 2
    u0 = op.initialize_dof_vector()
    op.project_initial(u0, 0.0)
    # eager: n1 + 1 + n2 steps, with the stage times of the calls below
    # (t0 + s dt from each call's start time; the Dirichlet data and the
    # speed factor follow the time, so its rounding must agree too)
    ue, r1, r2 = vectors(op, u0)
    run_eager(op, integ, ue, r1, r2, 0.0, dt, n1 + 1)
    run_eager(op, integ, ue, r1, r2, (n1 + 1) * dt, dt, n2)
    # a call of n1 steps (captures; the measured form times both), one
    # eager step, a call of n2 more (replays the cached graph)
    ug, g1, g2 = vectors(op, u0)
    integ.perform_time_steps(op, 0.0, dt, n1, ug, g1, g2)
    torch.cuda.synchronize()
    mid = ug.clone()
    integ.perform_time_step(op, n1 * dt, dt, ug, g1, g2)
    integ.perform_time_steps(op, (n1 + 1) * dt, dt, n2, ug, g1, g2)
    torch.cuda.synchronize()
    assert torch.equal(ug, ue), name
    ue1, e1, e2 = vectors(op, u0)
    run_eager(op, integ, ue1, e1, e2, 0.0, dt, n1)
    assert torch.equal(mid, ue1), name
    assert not torch.equal(ug, u0)
    info = op.step_launch()
    if mode:
        assert info["mode"] == mode
    elif name in AUTONOMOUS:
        # measured on the first call: both times recorded, the faster kept
        assert info["eager_ms"] > 0 and info["graph_ms"] > 0
        assert info["mode"] == (2 if info["graph_ms"] < info["eager_ms"]
                                else 1)
    else:
        assert info["mode"] == 0 and info["eager_ms"] == 0


@pytest.mark.parametrize("name", ["3d_k3_sipg_pert", "3d_k4_pert_cellular"])
def test_time_steps_as_first_call_of_an_operator(gpu_lib, name):
    """An operator whose first call is a multi-step call with graph replay
    forced: the tables built on first use (geometry, the SIPG face table)
    are built by the call's eager first step, before any capture; the
    result equals single steps on a second operator of the same
    configuration, bit for bit."""
    c = Cfg(**CASES[name])
    print(f"case {name}: first call with graph replay", flush=True)
    integ = LowStorageRungeKuttaIntegrator(1)
    dt, n = 1e-3, 2 * G + 3
    opg = AdvectionOperation(c.dg())
    opg.set_step_launch(2)
    ug = opg.initialize_dof_vector()
    opg.project_initial(ug, 0.0)
    u0 = ug.clone()
    g1, g2 = opg.initialize_dof_vector(), opg.initialize_dof_vector()
    integ.perform_time_steps(opg, 0.0, dt, n, ug, g1, g2)
    torch.cuda.synchronize()
    ope = AdvectionOperation(c.dg())
    ue, e1, e2 = vectors(ope, u0)
    run_eager(ope, integ, ue, e1, e2, 0.0, dt, n)
    torch.cuda.synchronize()
    assert opg.step_launch()["mode"] == 2
    assert not torch.equal(ug, u0)
    assert torch.equal(ug, ue), name


def test_graph_replays_between_eager_work(gpu_lib):
    """Several replays of one cached graph with eager steps and host
    synchronisations between them, each checked against the eager run."""
    c = Cfg(**CASES["3d_k4_pert_cellular"])
    op = AdvectionOperation(c.dg())
    op.set_step_launch(2)
    integ = LowStorageRungeKuttaIntegrator(3)
    dt = 2e-3
    u0 = op.initialize_dof_vector()
    op.project_initial(u0, 0.0)
    ug, g1, g2 = vectors(op, u0)
    ue, e1, e2 = vectors(op, u0)
    t = 0.0
    for rep in range(4):
        integ.perform_time_steps(op, t, dt, G + 1, ug, g1, g2)
        run_eager(op, integ, ue, e1, e2, t, dt, G + 1)
        t += (G + 1) * dt
        torch.cuda.synchronize()
        assert torch.equal(ug, ue), rep
        integ.perform_time_step(op, t, dt, ug, g1, g2)
        integ.perform_time_step(op, t, dt, ue, e1, e2)
        t += dt


def test_time_steps_edge_cases(gpu_lib):
    c = Cfg(**CASES["2d_k3_config1"])
    op = AdvectionOperation(c.dg())
    op.set_step_launch(2)
    integ = LowStorageRungeKuttaIntegrator(1)
    u = op.initialize_dof_vector()
    op.project_initial(u, 0.0)
    u0 = u.clone()
    r1, r2 = op.initialize_dof_vector(), op.initialize_dof_vector()
    integ.perform_time_steps(op, 0.0, 1e-3, 0, u, r1, r2)
    assert torch.equal(u, u0)
    # a different time step re-captures; other registers too
    n = G + 2
    integ.perform_time_steps(op, 0.0, 1e-3, n, u, r1, r2)
    integ.perform_time_steps(op, n * 1e-3, 2e-3, n, u, r1, r2)
    s1, s2 = op.initialize_dof_vector(), op.initialize_dof_vector()
    integ.perform_time_steps(op, n * 3e-3, 2e-3, n, u, s1, s2)
    ue, e1, e2 = vectors(op, u0)
    run_eager(op, integ, ue, e1, e2, 0.0, 1e-3, n)
    run_eager(op, integ, ue, e1, e2, n * 1e-3, 2e-3, 2 * n)
    torch.cuda.synchronize()
    assert torch.equal(u, ue)
    with pytest.raises(L.DgError):
        integ.perform_time_steps(op, 0.0, 1e-3, -1, u, r1, r2)
    with pytest.raises(L.DgError):
        op.set_step_launch(3)
