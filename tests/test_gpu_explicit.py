"""GPU parity of the explicit operator, the fused LSRK stage and the
auxiliary integrals against the CPU oracle (tolerance 1e-12 relative in
FP64, index data bit-exact)."""
import numpy as np
import pytest
import torch

from advection_miniapp_amd import AdvectionOperation, \
    LowStorageRungeKuttaIntegrator
from helpers import (Cfg, nodal_points, rel_diff, transport_error,
                     transport_problem)

pytestmark = pytest.mark.gpu

TOL = 1e-12

CASES = {
    # name: config
    "2d_k3_cart_const": dict(dim=2, degree=3, refine=3, lo=[-1, -1],
                             hi=[1, 1], velocity=[1.0, 0.15]),
    "2d_k4_pert_vortex": dict(dim=2, degree=4, refine=3, perturbation=0.1,
                              velocity_kind=2, time_factor=1,
                              final_time=8.0, flux_alpha=0.0),
    "2d_k2_dirichlet": dict(dim=2, degree=2, n_coarse=[3, 2, 1], refine=2,
                            periodic=[0, 0, 0], velocity=[1.1, 0.15],
                            solution_kind=1),
    "3d_k4_cart_const": dict(dim=3, degree=4, refine=2,
                             velocity=[1.0, 0.5, 0.25]),
    "3d_k4_pert_cellular": dict(dim=3, degree=4, refine=2, perturbation=0.1,
                                velocity_kind=1),
    "3d_k1_pert": dict(dim=3, degree=1, refine=3, perturbation=0.2,
                       velocity_kind=1),
    "3d_k2_dirichlet_pert": dict(dim=3, degree=2, n_coarse=[2, 1, 1],
                                 refine=1, periodic=[0, 1, 0],
                                 perturbation=0.15,
                                 velocity=[1.1, 0.15, -0.05],
                                 solution_kind=1),
    "3d_k3_coarse3": dict(dim=3, degree=3, n_coarse=[3, 3, 3], refine=1,
                          lo=[-1, -1, -1], hi=[1, 1, 1],
                          velocity=[0.3, -0.2, 0.9]),
    "3d_k5_pert": dict(dim=3, degree=5, refine=1, perturbation=0.1,
                       velocity_kind=1),
    "3d_k6_cart": dict(dim=3, degree=6, refine=1, velocity=[1, 0.5, 0.25]),
    "3d_k7_cart": dict(dim=3, degree=7, refine=1, velocity=[-1, 0.5, 0.25]),
    "2d_k7_pert": dict(dim=2, degree=7, refine=2, perturbation=0.1,
                       velocity_kind=3, time_factor=1, final_time=8.0),
    # skew-symmetric form of the explicit operator (factor_skew,
    # advection_solver.cc:106-108)
    "3d_k3_pert_skew": dict(dim=3, degree=3, refine=2, perturbation=0.1,
                            velocity_kind=1, factor_skew=0.5),
    "2d_k4_vortex_skew": dict(dim=2, degree=4, refine=3, perturbation=0.1,
                              velocity_kind=2, time_factor=1,
                              final_time=8.0, factor_skew=0.3,
                              flux_alpha=0.0),
    "3d_k2_dirichlet_skew": dict(dim=3, degree=2, n_coarse=[2, 1, 1],
                                 refine=1, periodic=[0, 1, 0],
                                 velocity=[1.1, 0.15, -0.05],
                                 solution_kind=1, factor_skew=0.5),
}


def make(name, **extra):
    kw = dict(CASES[name])
    kw.update(extra)
    c = Cfg(**kw)
    return c, AdvectionOperation(c.dg()), c.oracle_handle()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("name", sorted(CASES))
def test_index_data_bit_exact(gpu_lib, name):
    c, op, o = make(name)
    assert np.array_equal(op.cell_coordinates(), o.cell_coordinates())
    assert np.array_equal(op.neighbors(), o.neighbors())
    assert np.array_equal(op.cell_vertices(), o.cell_vertices())
    assert op.min_vertex_distance == o.min_vertex_distance()


@pytest.mark.parametrize("name", sorted(CASES))
def test_apply_matches_oracle(gpu_lib, name):
    c, op, o = make(name)
    u = np.random.default_rng(11).standard_normal(o.n_dofs)
    for t in (0.0, 0.37):
        ref = o.apply(u, t)
        dst = op.initialize_dof_vector()
        op.apply(dev(u), dst, t)
        torch.cuda.synchronize()
        assert rel_diff(host(dst), ref) < TOL, name


def test_apply_known_answer_polynomial(gpu_lib):
    """u in Q_k continuous with Dirichlet data u: M^-1(-Lu) = -a.grad u."""
    k, dim = 4, 3
    a = [0.7, -0.4, 0.3]
    c = Cfg(dim=dim, degree=k, refine=1, periodic=[0, 0, 0], velocity=a,
            solution_kind=4, solution_param=[k], lo=[-0.5, 0, 0.25],
            hi=[0.5, 1.5, 1.0])
    op = AdvectionOperation(c.dg())
    u = op.initialize_dof_vector()
    op.project_initial(u, 0.0)
    du = op.initialize_dof_vector()
    op.apply(u, du, 0.0)
    pts = nodal_points(op.cell_vertices(), k + 1, dim)
    s = 1 + sum((d + 1) * pts[..., d] for d in range(dim))
    exact = -k * s ** (k - 1) * sum((d + 1) * a[d] for d in range(dim))
    assert rel_diff(host(du), exact.ravel()) < 1e-11
    assert rel_diff(host(u), (s ** k).ravel()) < 1e-12


@pytest.mark.parametrize("dim,k", [(3, 2), (3, 4), (2, 3)])
def test_apply_known_answer_linear_on_perturbed_cells(gpu_lib, dim, k):
    """Non-affine cells: a globally linear u lies in the mapped Q_k space
    and QGauss(k+1) integrates every term exactly (k >= 2 in 3D), so
    M^-1(-L u) = -a.grad u, a constant; a transposed Jacobian contraction
    misses it by a factor of ~100 (the oracle's pin of the same case)."""
    a = [0.7, -0.4, 0.3][:dim] + [0.0] * (3 - dim)
    c = Cfg(dim=dim, degree=k, refine=2, periodic=[0, 0, 0],
            perturbation=0.15, seed=11, velocity=a, solution_kind=4,
            solution_param=[1], lo=[-0.5, 0, 0.25], hi=[0.5, 1.5, 1.0])
    op = AdvectionOperation(c.dg())
    u = op.initialize_dof_vector()
    op.project_initial(u, 0.0)
    du = op.initialize_dof_vector()
    op.apply(u, du, 0.0)
    grad_dot = sum((d + 1) * a[d] for d in range(dim))
    assert rel_diff(host(du), np.full(op.n_dofs, -grad_dot)) < 1e-10


@pytest.mark.parametrize("velocity_kind", [0, 1])
def test_apply_perturbed_mesh_transport_converges(gpu_lib, velocity_kind):
    """The HIP apply at degree 4 on 3D cells perturbed by 10% of h:
    M^-1(-L u) of u = prod sin(2 pi x_d) approaches -a.grad u at about
    order k over refine 2 -> 4 (64 -> 4096 cells), for a constant speed and
    for the cellular field."""
    k, errs = 4, []
    for r in (2, 3, 4):
        c = transport_problem(k, r, velocity_kind)
        op = AdvectionOperation(c.dg())
        u = op.initialize_dof_vector()
        op.project_initial(u, 0.0)
        du = op.initialize_dof_vector()
        op.apply(u, du, 0.0)
        errs.append(transport_error(host(du), op.cell_vertices(), k,
                                    velocity_kind))
        del op, u, du
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > k - 0.75), (errs, rates)
    assert errs[-1] < 1e-4, errs


@pytest.mark.parametrize("name", ["3d_k4_pert_cellular", "2d_k4_pert_vortex",
                                  "2d_k2_dirichlet", "3d_k3_pert_skew",
                                  "3d_k2_dirichlet_skew"])
def test_perform_stage_matches_oracle(gpu_lib, name):
    c, op, o = make(name)
    rng = np.random.default_rng(3)
    ri, sol = rng.standard_normal((2, o.n_dofs))
    t, b, a = 0.2, 0.013, 0.021
    # separate registers: fused kernel
    ki_o, sol_o, nri_o = np.zeros(o.n_dofs), sol.copy(), np.zeros(o.n_dofs)
    o.perform_stage(t, b, a, ri, ki_o, sol_o, nri_o)
    d_ri, d_sol = dev(ri), dev(sol)
    d_ki, d_nri = op.initialize_dof_vector(), op.initialize_dof_vector()
    op.perform_stage(t, b, a, d_ri, d_ki, d_sol, d_nri)
    assert rel_diff(host(d_sol), sol_o) < TOL
    assert rel_diff(host(d_nri), nri_o) < TOL
    # aliased registers as in the reference integrator's first stage
    sol_o2, r_o2 = sol.copy(), np.zeros(o.n_dofs)
    o.perform_stage(t, b, a, sol_o2, r_o2, sol_o2, r_o2)
    d_sol2, d_r2 = dev(sol), op.initialize_dof_vector()
    op.perform_stage(t, b, a, d_sol2, d_r2, d_sol2, d_r2)
    assert rel_diff(host(d_sol2), sol_o2) < TOL
    assert rel_diff(host(d_r2), r_o2) < TOL
    # last stage (a = 0)
    sol_o3, ki3, r3 = sol.copy(), np.zeros(o.n_dofs), ri.copy()
    o.perform_stage(t, b, 0.0, r3, ki3, sol_o3, r3)
    d_sol3, d_r3 = dev(sol), dev(ri)
    d_ki3 = op.initialize_dof_vector()
    op.perform_stage(t, b, 0.0, d_r3, d_ki3, d_sol3, d_r3)
    assert rel_diff(host(d_sol3), sol_o3) < TOL


@pytest.mark.parametrize("scheme", [0, 1, 2, 3])
def test_lsrk_time_steps_match_oracle(gpu_lib, scheme):
    c, op, o = make("3d_k3_coarse3", solution_kind=1)
    u = o.project(0.0)
    r1, r2 = o.zeros(), o.zeros()
    d_u = dev(u)
    d_r1, d_r2 = op.initialize_dof_vector(), op.initialize_dof_vector()
    integ = LowStorageRungeKuttaIntegrator(scheme)
    t, dt = 0.0, 0.01
    for _ in range(3):
        o.lsrk_time_step(scheme, t, dt, u, r1, r2)
        integ.perform_time_step(op, t, dt, d_u, d_r1, d_r2)
        t += dt
    assert rel_diff(host(d_u), u) < TOL


@pytest.mark.parametrize("name", ["3d_k4_pert_cellular", "2d_k3_cart_const",
                                  "3d_k2_dirichlet_pert"])
def test_projection_and_integrals_match_oracle(gpu_lib, name):
    c, op, o = make(name, solution_kind=2, solution_param=[2 * np.pi, 0])
    ref = o.project(0.0)
    u = op.initialize_dof_vector()
    op.project_initial(u, 0.0)
    assert rel_diff(host(u), ref) < TOL
    assert np.allclose(op.compute_mass_and_energy(u), o.mass_energy(ref),
                       rtol=1e-12, atol=1e-14)
    uu = dev(ref + 0.01 * np.sin(np.arange(ref.size)))
    assert np.allclose(op.l2_error(uu, 0.1), o.l2_error(host(uu), 0.1),
                       rtol=1e-12, atol=0)


@pytest.mark.parametrize("name", ["3d_k4_pert_cellular", "3d_k4_cart_const",
                                  "2d_k3_cart_const", "2d_k2_dirichlet"])
def test_fp32_apply_error_against_fp64_oracle(gpu_lib, name):
    """FP32 operator (tabulated and uniform geometry): the error stays
    within a small multiple of the FP32 rounding of the input pushed
    through the FP64 operator."""
    c = Cfg(**CASES[name])
    o = c.oracle_handle()
    c.precision = 1
    op = AdvectionOperation(c.dg())
    u = np.random.default_rng(8).standard_normal(o.n_dofs)
    dst = op.initialize_dof_vector()
    op.apply(dev(u).float(), dst, 0.1)
    ref = o.apply(u, 0.1)
    floor = rel_diff(o.apply(u.astype(np.float32).astype(np.float64), 0.1),
                     ref)
    err = rel_diff(dst.double().cpu().numpy(), ref)
    assert err < max(20 * floor, 2e-6), (err, floor)


def test_errors_are_reported(gpu_lib):
    c, op, o = make("3d_k4_cart_const")
    u = op.initialize_dof_vector()
    with pytest.raises(ValueError):
        op.apply(u, torch.zeros(5, dtype=torch.float64, device="cuda"), 0.0)
    with pytest.raises(TypeError):
        op.apply(u, u.float(), 0.0)
    with pytest.raises(ValueError):
        op.apply(u.cpu(), u, 0.0)
