"""GPU parity of the advection + SIPG diffusion operator (explicit right-hand
side, fused LSRK stage, implicit operator) against the CPU oracle; FP32
operator error against the FP64 oracle."""
import numpy as np
import pytest
import torch

from advection_miniapp_amd import AdvectionOperation
from helpers import Cfg, rel_diff

pytestmark = pytest.mark.gpu


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def host(t):
    return t.detach().double().cpu().numpy()


def sipg(dim, k, **kw):
    base = dict(dim=dim, degree=k, refine=2 if dim == 3 else 3,
                velocity=[1.0, 0.5, 0.25], operator_kind=1, viscosity=1e-2,
                penalty_factor=(k + 1) ** 2, solution_kind=2,
                solution_param=[2 * np.pi, 1e-2])
    base.update(kw)
    return Cfg(**base)


@pytest.mark.parametrize("dim,k", [(3, 3), (3, 5), (3, 7), (2, 4), (2, 2)])
def test_sipg_apply_matches_oracle(gpu_lib, dim, k):
    c = sipg(dim, k, refine=1 if k == 7 else (2 if dim == 3 else 3))
    op, o = AdvectionOperation(c.dg()), c.oracle_handle()
    u = np.random.default_rng(12).standard_normal(o.n_dofs)
    dst = op.initialize_dof_vector()
    op.apply(dev(u), dst, 0.1)
    assert rel_diff(host(dst), o.apply(u, 0.1)) < 1e-12


@pytest.mark.parametrize("dim,k,vk", [(3, 4, 1), (2, 3, 2)])
def test_sipg_variable_speed_matches_oracle(gpu_lib, dim, k, vk):
    """A space-dependent speed takes the per-point speed evaluation (the
    constant-speed cases above use products of 1D weights)."""
    c = sipg(dim, k, velocity_kind=vk)
    op, o = AdvectionOperation(c.dg()), c.oracle_handle()
    u = np.random.default_rng(13).standard_normal(o.n_dofs)
    dst = op.initialize_dof_vector()
    op.apply(dev(u), dst, 0.3)
    assert rel_diff(host(dst), o.apply(u, 0.3)) < 1e-12
    op.set_time(0.2, 0.05)
    dst = op.initialize_dof_vector()
    op.vmult(dst, dev(u))
    assert rel_diff(host(dst), o.vmult(u, 0.2, 0.05)) < 1e-12


@pytest.mark.parametrize("k", [3, 5, 7])
def test_sipg_fp32_error(gpu_lib, k):
    c = sipg(3, k, refine=1 if k == 7 else 2)
    o = c.oracle_handle()
    c.precision = 1
    op = AdvectionOperation(c.dg())
    u = o.project(0.0)
    dst = op.initialize_dof_vector()
    op.apply(dev(u, torch.float32), dst, 0.0)
    err = rel_diff(host(dst), o.apply(u, 0.0))
    # FP32 rounding of the input alone, pushed through the FP64 operator:
    # the second-derivative terms amplify it by ~k^4 / h^2 / |L u|
    u32 = u.astype(np.float32).astype(np.float64)
    floor = rel_diff(o.apply(u32, 0.0), o.apply(u, 0.0))
    print(f"sipg fp32 k {k}: error {err:.2e}, input rounding floor "
          f"{floor:.2e}")
    assert err < max(20 * floor, 2e-5), (err, floor)


def test_sipg_lsrk_stage_matches_oracle(gpu_lib):
    c = sipg(3, 3)
    op, o = AdvectionOperation(c.dg()), c.oracle_handle()
    rng = np.random.default_rng(13)
    ri, sol = rng.standard_normal((2, o.n_dofs))
    ki_o, sol_o, nri_o = np.zeros(o.n_dofs), sol.copy(), np.zeros(o.n_dofs)
    o.perform_stage(0.1, 0.01, 0.02, ri, ki_o, sol_o, nri_o)
    d_sol, d_ki, d_nri = dev(sol), op.initialize_dof_vector(), \
        op.initialize_dof_vector()
    op.perform_stage(0.1, 0.01, 0.02, dev(ri), d_ki, d_sol, d_nri)
    assert rel_diff(host(d_sol), sol_o) < 1e-12
    assert rel_diff(host(d_nri), nri_o) < 1e-12


def test_sipg_implicit_vmult_matches_oracle(gpu_lib):
    c = sipg(3, 4, factor_skew=0.5, penalty_factor=2 * 4 * 5)
    op, o = AdvectionOperation(c.dg()), c.oracle_handle()
    u = np.random.default_rng(14).standard_normal(o.n_dofs)
    op.set_time(0.2, 0.05)
    dst = op.initialize_dof_vector()
    op.vmult(dst, dev(u))
    assert rel_diff(host(dst), o.vmult(u, 0.2, 0.05)) < 1e-12


# SIPG on perturbed (non-axis-aligned) cells: advection operator kernel in
# integrate form + the general-geometry diffusion kernel + inverse mass
PERTURBED = {
    "3d_k3_const": dict(dim=3, k=3, perturbation=0.1, viscosity=0.05),
    "3d_k4_cellular": dict(dim=3, k=4, perturbation=0.15, viscosity=0.02,
                           velocity_kind=1),
    "3d_k2_pure_diffusion": dict(dim=3, k=2, perturbation=0.1,
                                 viscosity=1.0, velocity=[0, 0, 0]),
    "3d_k7_const": dict(dim=3, k=7, perturbation=0.1, viscosity=0.01,
                        refine=1),
    "2d_k4_vortex": dict(dim=2, k=4, perturbation=0.1, viscosity=0.05,
                         velocity_kind=2, time_factor=1, final_time=8.0),
    "2d_k1": dict(dim=2, k=1, perturbation=0.2, viscosity=0.1),
}


def sipg_perturbed(name):
    kw = dict(PERTURBED[name])
    dim, k = kw.pop("dim"), kw.pop("k")
    kw.setdefault("seed", 7)
    return sipg(dim, k, **kw)


@pytest.mark.parametrize("name", sorted(PERTURBED))
def test_sipg_perturbed_apply_matches_oracle(gpu_lib, name):
    c = sipg_perturbed(name)
    print(f"case {name}", flush=True)
    op, o = AdvectionOperation(c.dg()), c.oracle_handle()
    u = np.random.default_rng(15).standard_normal(o.n_dofs)
    for t in (0.0, 0.3):
        dst = op.initialize_dof_vector()
        op.apply(dev(u), dst, t)
        assert rel_diff(host(dst), o.apply(u, t)) < 1e-12, (name, t)


@pytest.mark.parametrize("name", ["3d_k3_const", "2d_k4_vortex"])
def test_sipg_perturbed_stage_and_vmult_match_oracle(gpu_lib, name):
    c = sipg_perturbed(name)
    op, o = AdvectionOperation(c.dg()), c.oracle_handle()
    rng = np.random.default_rng(16)
    ri, sol = rng.standard_normal((2, o.n_dofs))
    ki_o, sol_o, nri_o = np.zeros(o.n_dofs), sol.copy(), np.zeros(o.n_dofs)
    o.perform_stage(0.1, 0.01, 0.02, ri, ki_o, sol_o, nri_o)
    d_sol, d_ki, d_nri = dev(sol), op.initialize_dof_vector(), \
        op.initialize_dof_vector()
    op.perform_stage(0.1, 0.01, 0.02, dev(ri), d_ki, d_sol, d_nri)
    assert rel_diff(host(d_sol), sol_o) < 1e-12
    assert rel_diff(host(d_nri), nri_o) < 1e-12
    # implicit operator M/dt + L(t + dt)
    op.set_time(0.2, 0.05)
    dst = op.initialize_dof_vector()
    op.vmult(dst, dev(ri))
    assert rel_diff(host(dst), o.vmult(ri, 0.2, 0.05)) < 1e-12


def test_sipg_perturbed_mesh_converges(gpu_lib):
    """Pure diffusion (nu = 1) at degree 4 on cells perturbed by 10% of h:
    the HIP M^-1(-L u) of u = prod sin(2 pi x_d) approaches nu lap u at
    about order k - 1 (the oracle's pin of the same case)."""
    from helpers import nodal_points
    k, errs = 4, []
    for r in (1, 2, 3):
        c = Cfg(dim=3, degree=k, refine=r, perturbation=0.1, seed=7,
                velocity=[0, 0, 0], operator_kind=1, viscosity=1.0,
                penalty_factor=2 * k * (k + 1), solution_kind=2,
                solution_param=[2 * np.pi, 0.0])
        op = AdvectionOperation(c.dg())
        u = op.initialize_dof_vector()
        op.project_initial(u, 0.0)
        du = op.initialize_dof_vector()
        op.apply(u, du, 0.0)
        p = nodal_points(op.cell_vertices(), k + 1, 3)
        ex = (-3 * (2 * np.pi) ** 2 * np.prod(np.sin(2 * np.pi * p), -1))
        ex = ex.ravel()
        d = host(du)
        errs.append(np.sqrt(np.mean((d - ex) ** 2) / np.mean(ex ** 2)))
        del op, u, du
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > k - 2), (errs, rates)
    assert errs[-1] < 3e-3, errs


def test_sipg_perturbed_fp32_and_irk_are_refused(gpu_lib):
    from advection_miniapp_amd import _lib as L
    c = sipg_perturbed("3d_k3_const")
    c.precision = 1
    op = AdvectionOperation(c.dg())
    dst = op.initialize_dof_vector()
    with pytest.raises(L.DgError) as e:
        op.apply(op.initialize_dof_vector(), dst, 0.0)
    assert e.value.code == L.DG_ERR_UNSUPPORTED
    # the Radau IIA operators are built for advection, as the reference's
    # IRK solvers (on perturbed and on axis-aligned cells)
    for cfg in (sipg_perturbed("3d_k3_const"), sipg(3, 3)):
        op = AdvectionOperation(cfg.dg())
        op.set_time(0.0, 0.01)
        s = cfg.irk_stages
        z = op.initialize_dof_vector(blocks=s)
        with pytest.raises(L.DgError) as e:
            op.vmult_irk(op.initialize_dof_vector(blocks=s), z)
        assert e.value.code == L.DG_ERR_UNSUPPORTED
        with pytest.raises(L.DgError) as e:
            op.compute_rhs(z, op.initialize_dof_vector())
        assert e.value.code == L.DG_ERR_UNSUPPORTED
