"""Seeded random problem configurations: the GPU apply, fused LSRK stage,
implicit and Radau IIA operators and the auxiliary integrals against the
oracle at 1e-12 relative (FP64). Each draw picks the dimension,
degree, coarse mesh, refinement, box, periodicity per direction,
perturbation, speed field, time factor, flux and skew factors, operator
kind (advection, SIPG on periodic meshes, Burgers) and the quadrature
switches, within what the library accepts, so that combinations the named
cases of the other tests do not list are covered too. The draw is printed
before it runs."""
import numpy as np
import pytest
import torch

from advection_miniapp_amd import AdvectionOperation
from advection_miniapp_amd import _lib as L
from helpers import Cfg, rel_diff

pytestmark = pytest.mark.gpu

TOL = 1e-12
MAX_DOFS = 40000


def draw(seed):
    rng = np.random.default_rng(1000 + seed)
    dim = int(rng.choice([2, 3]))
    kind = str(rng.choice(["advection"] * 3 + ["sipg", "burgers"]))
    kmax = 4 if (kind == "burgers" and dim == 2) else 7
    k = int(rng.integers(1, kmax + 1))
    nc = [int(x) for x in rng.integers(1, 4, size=dim)]
    ncomp = dim if kind == "burgers" else 1
    nd = (k + 1) ** dim * ncomp
    if np.prod(nc) * 2 ** dim * nd > MAX_DOFS:
        nc = [1] * dim
    refine = 1
    while (np.prod(nc) * 2 ** (dim * (refine + 1)) * nd <= MAX_DOFS
           and refine < 4):
        refine += 1
    refine = int(rng.integers(1, refine + 1))
    lo = [float(x) for x in rng.uniform(-1.0, 0.0, size=dim)]
    hi = [lo[d] + float(rng.uniform(0.5, 2.0)) for d in range(dim)]
    periodic = [1] * dim
    if kind == "advection":
        periodic = [int(x) for x in rng.integers(0, 2, size=dim)]
    kw = dict(dim=dim, degree=k, n_coarse=nc, refine=refine, lo=lo, hi=hi,
              periodic=periodic,
              perturbation=float(rng.choice([0.0, 0.05, 0.15, 0.25])),
              seed=int(rng.integers(1, 1000)))
    if kind == "burgers":
        kw.update(operator_kind=2, n_components=dim, solution_kind=6,
                  flux_alpha=float(rng.choice([0.5, 1.0])))
        return kw
    vk = int(rng.choice([0, 0, 2, 3] + ([1] if dim == 3 else [])))
    if vk != 0:
        # the analytic fields are 1-periodic and live on the unit box (the
        # reference's vortex and cellular problems); on other boxes they are
        # discontinuous across periodic faces, where the face speed then
        # depends on the side that evaluates it
        kw.update(lo=[0.0] * dim, hi=[1.0] * dim)
    kw.update(velocity_kind=vk,
              velocity=[float(x) for x in rng.uniform(-1.2, 1.2, size=3)],
              flux_alpha=float(rng.choice([0.0, 0.5, 1.0])),
              solution_kind=int(rng.choice([1, 2])),
              solution_param=[np.pi, 0.0])
    if vk in (2, 3) and rng.random() < 0.7:
        kw.update(time_factor=1, final_time=float(rng.uniform(2.0, 8.0)))
    if kind == "sipg":
        kw.update(operator_kind=1,
                  viscosity=float(rng.choice([1e-3, 1e-2, 0.1])),
                  penalty_factor=float((k + 1) ** 2))
    else:
        kw.update(factor_skew=float(rng.choice([0.0, 0.0, 0.5, 1.0])))
        g = rng.random()
        if g < 0.15:
            kw.update(use_gl_quad=1)
        elif g < 0.3:
            kw.update(use_gl_quad_mass=1)
    return kw


SEEDS = list(range(64))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def test_draws_cover_the_axes():
    """The draws (no GPU work) span the configuration axes they claim."""
    ds = [draw(s) for s in SEEDS]
    assert {d["dim"] for d in ds} == {2, 3}
    assert {d["degree"] for d in ds} >= {1, 2, 3, 4, 5, 6, 7}
    assert {d.get("operator_kind", 0) for d in ds} == {0, 1, 2}
    assert any(0 in d["periodic"] for d in ds)
    assert any(d["perturbation"] > 0 for d in ds)
    assert any(d.get("time_factor", 0) == 1 for d in ds)
    assert any(d.get("factor_skew", 0) > 0 for d in ds)
    assert any(d.get("use_gl_quad", 0) or d.get("use_gl_quad_mass", 0)
               for d in ds)


@pytest.mark.parametrize("seed", SEEDS)
def test_random_config_apply_and_stage(gpu_lib, seed):
    kw = draw(seed)
    print(f"seed {seed}: {kw}", flush=True)
    c = Cfg(**kw)
    op, o = AdvectionOperation(c.dg()), c.oracle_handle()
    # mesh and index data bit-exact
    assert np.array_equal(op.cell_coordinates(), o.cell_coordinates())
    assert np.array_equal(op.neighbors(), o.neighbors())
    assert np.array_equal(op.cell_vertices(), o.cell_vertices())
    assert op.min_vertex_distance == o.min_vertex_distance()
    rng = np.random.default_rng(seed)
    n = o.n_dofs * c.n_components
    u = rng.standard_normal(n)
    burgers = kw.get("operator_kind", 0) == 2
    for t in (0.0, 0.37):
        dst = op.initialize_dof_vector()
        op.apply(dev(u), dst, t)
        ref = o.burgers_apply(u) if burgers else o.apply(u, t)
        assert rel_diff(host(dst), ref) < TOL, (seed, t)
        if burgers:
            break
    # one fused stage with distinct registers (the oracle's stage runs the
    # scalar operator; the Burgers stage is formed from its right-hand side)
    ri, sol = rng.standard_normal((2, n))
    ki_o, sol_o, nri_o = np.zeros(n), sol.copy(), np.zeros(n)
    if burgers:
        k_o = o.burgers_apply(ri)
        sol_o, nri_o = sol + 0.013 * k_o, sol + 0.029 * k_o
    else:
        o.perform_stage(0.21, 0.013, 0.029, ri, ki_o, sol_o, nri_o)
    d_sol, d_ki, d_nri = dev(sol), op.initialize_dof_vector(), \
        op.initialize_dof_vector()
    op.perform_stage(0.21, 0.013, 0.029, dev(ri), d_ki, d_sol, d_nri)
    assert rel_diff(host(d_sol), sol_o) < TOL, seed
    assert rel_diff(host(d_nri), nri_o) < TOL, seed


def implicit_draw(kw):
    """The implicit operator is built for advection and SIPG with Gauss
    quadrature (not for Burgers or the quadrature switches)."""
    return (kw.get("operator_kind", 0) != 2 and not kw.get("use_gl_quad", 0)
            and not kw.get("use_gl_quad_mass", 0))


VMULT_SEEDS = [s for s in SEEDS if implicit_draw(draw(s))]


@pytest.mark.parametrize("seed", VMULT_SEEDS)
def test_random_config_implicit_vmult(gpu_lib, seed):
    """M/dt + L(t + dt) (and L alone for dt = 0) against the oracle."""
    kw = draw(seed)
    print(f"seed {seed}: {kw}", flush=True)
    c = Cfg(**kw)
    op, o = AdvectionOperation(c.dg()), c.oracle_handle()
    u = np.random.default_rng(50 + seed).standard_normal(o.n_dofs)
    for t, dt in ((0.1, 0.02), (0.3, 0.0)):
        op.set_time(t, dt)
        dst = op.initialize_dof_vector()
        op.vmult(dst, dev(u))
        assert rel_diff(host(dst), o.vmult(u, t, dt)) < TOL, (seed, t, dt)


def irk_draw(kw):
    """The Radau IIA operators: advection (as the reference's IRK solvers),
    Gauss quadrature, up to degree 6."""
    return (implicit_draw(kw) and kw["degree"] <= 6
            and kw.get("operator_kind", 0) == 0)


IRK_SEEDS = [s for s in SEEDS if irk_draw(draw(s))]


@pytest.mark.parametrize("seed", IRK_SEEDS)
def test_random_config_irk_operators(gpu_lib, seed):
    """The stage-coupled operator, its right-hand side and the solution
    update of Radau IIA with 2 or 3 stages (the 4-stage table of the
    reference holds 7-digit literals) against the oracle."""
    kw = dict(draw(seed), irk_stages=2 + seed % 2)
    print(f"seed {seed}: {kw}", flush=True)
    c = Cfg(**kw)
    s = kw["irk_stages"]
    op, o = AdvectionOperation(c.dg()), c.oracle_handle()
    rng = np.random.default_rng(80 + seed)
    u = rng.standard_normal(o.n_dofs)
    Z = rng.standard_normal(o.n_dofs * s)
    t, dt = 0.2, 0.05
    op.set_time(t, dt)
    rhs = op.initialize_dof_vector(blocks=s)
    op.compute_rhs(rhs, dev(u))
    assert rel_diff(host(rhs), o.irk_compute_rhs(u, t, dt)) < TOL, seed
    dst = op.initialize_dof_vector(blocks=s)
    op.vmult_irk(dst, dev(Z))
    assert rel_diff(host(dst), o.irk_vmult(Z, t, dt)) < TOL, seed
    sol = dev(u)
    op.update_solution(sol, dev(Z))
    u_o = u.copy()
    o.irk_update_solution(u_o, Z, dt)
    assert rel_diff(host(sol), u_o) < TOL, seed


AUX_SEEDS = [s for s in SEEDS if draw(s).get("operator_kind", 0) != 2]


@pytest.mark.parametrize("seed", AUX_SEEDS)
def test_random_config_auxiliary(gpu_lib, seed):
    """Projection of the analytic solution (at two times), mass and
    energy, the L2 error and the standalone inverse mass (scalar fields)
    against the oracle."""
    kw = draw(seed)
    print(f"seed {seed}: {kw}", flush=True)
    c = Cfg(**kw)
    op, o = AdvectionOperation(c.dg()), c.oracle_handle()
    for t in (0.0, 0.37):
        ref = o.project(t)
        u = op.initialize_dof_vector()
        op.project_initial(u, t)
        assert rel_diff(host(u), ref) < TOL, (seed, t)
    assert np.allclose(op.compute_mass_and_energy(u), o.mass_energy(ref),
                       rtol=1e-12, atol=1e-14), seed
    w = ref + 0.01 * np.sin(np.arange(ref.size))
    assert np.allclose(op.l2_error(dev(w), 0.37), o.l2_error(w, 0.37),
                       rtol=1e-12, atol=0), seed
    r = np.random.default_rng(90 + seed).standard_normal(o.n_dofs)
    d = dev(r)
    op.apply_inverse_mass_matrix(d)
    assert rel_diff(host(d), o.inverse_mass(r)) < TOL, seed


def fp32_draw(kw):
    """FP32 operators: advection, and SIPG on axis-aligned cells."""
    k = kw.get("operator_kind", 0)
    return k == 0 or (k == 1 and kw["perturbation"] == 0)


FP32_SEEDS = [s for s in SEEDS if fp32_draw(draw(s))]


@pytest.mark.parametrize("seed", FP32_SEEDS)
def test_random_config_fp32_apply_error(gpu_lib, seed):
    """The FP32 apply stays within a small multiple of the error that
    rounding the input to FP32 causes through the FP64 operator."""
    kw = dict(draw(seed), precision=1)
    print(f"seed {seed}: {kw}", flush=True)
    c = Cfg(**kw)
    op = AdvectionOperation(c.dg())
    c.precision = 0
    o = c.oracle_handle()
    u = np.random.default_rng(70 + seed).standard_normal(o.n_dofs)
    dst = op.initialize_dof_vector()
    op.apply(dev(u).float(), dst, 0.1)
    ref = o.apply(u, 0.1)
    floor = rel_diff(o.apply(u.astype(np.float32).astype(np.float64), 0.1),
                     ref)
    err = rel_diff(host(dst), ref)
    assert err < max(20 * floor, 2e-6), (seed, err, floor)


@pytest.mark.parametrize("seed", SEEDS[::2])
def test_random_config_time_loop_in_one_call(gpu_lib, seed):
    """dg_lsrk_time_steps with graph replay forced equals the same steps
    taken one by one, bit for bit (time-dependent operators and boundary
    faces take the per-step path inside the call)."""
    from advection_miniapp_amd import LowStorageRungeKuttaIntegrator
    kw = draw(seed)
    print(f"seed {seed}: {kw}", flush=True)
    op = AdvectionOperation(Cfg(**kw).dg())
    op.set_step_launch(2)
    integ = LowStorageRungeKuttaIntegrator(L.DG_LSRK_5_4)
    u0 = op.initialize_dof_vector()
    op.project_initial(u0, 0.0)
    dt, n = 1e-3, 2 * L.DG_GRAPH_STEPS + 3
    ue, e1, e2 = u0.clone(), op.initialize_dof_vector(), \
        op.initialize_dof_vector()
    for s in range(n):
        integ.perform_time_step(op, s * dt, dt, ue, e1, e2)
    ug, g1, g2 = u0.clone(), op.initialize_dof_vector(), \
        op.initialize_dof_vector()
    integ.perform_time_steps(op, 0.0, dt, n, ug, g1, g2)
    torch.cuda.synchronize()
    assert torch.equal(ug, ue), seed
