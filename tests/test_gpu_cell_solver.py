"""The batched cell solver of the IRK block-Jacobi preconditioner
(do_batched_solver, advection_solver_irk_real.cc:2079-2126 and the
CellwiseOperator of :525-782): the cell-local stage operator in the
Gauss-point collocation basis against the oracle, the exactness of the
preconditioner where the cell operator is separable, the minimal-residual
property of its GMRES step, and IRK solves with it."""
import numpy as np
import pytest
import torch

from advection_miniapp_amd import AdvectionOperation
from advection_miniapp_amd import _lib as L
from advection_miniapp_amd.operators import solver_control
from helpers import Cfg

pytestmark = pytest.mark.gpu


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def host(t):
    return t.detach().double().cpu().numpy()


def basis_change(k):
    """S[q][i] = Gauss-Lobatto Lagrange basis i at Gauss point q on [0,1]
    (numpy's Legendre roots, independent of the library)."""
    from numpy.polynomial import legendre
    n = k + 1
    inner = legendre.Legendre.basis(k).deriv().roots() if k > 1 else []
    xl = (np.concatenate(([-1.0], np.sort(inner), [1.0])) + 1) / 2
    xg = (np.sort(legendre.leggauss(n)[0]) + 1) / 2
    S = np.ones((n, n))
    for i in range(n):
        for m in range(n):
            if m != i:
                S[:, i] *= (xg - xl[m]) / (xl[i] - xl[m])
    return S


def apply3(M, u, n, dim):
    """M applied along every direction of cell-wise lexicographic data
    u[..., cell, nd] (x fastest)."""
    sh = u.shape[:-1]
    if dim == 3:
        v = u.reshape(*sh, n, n, n)
        v = np.einsum("ai,...kji->...kja", M, v)
        v = np.einsum("bj,...kja->...kba", M, v)
        v = np.einsum("ck,...kba->...cba", M, v)
    else:
        v = u.reshape(*sh, n, n)
        v = np.einsum("ai,...ji->...ja", M, v)
        v = np.einsum("bj,...ja->...ba", M, v)
    return v.reshape(u.shape)


def cell_parity(o, c):
    """(i + j + k) mod 2 of every cell from its centre."""
    V = o.cell_vertices().reshape(o.n_cells, 2 ** c.dim, c.dim)
    ctr = V.mean(axis=1)
    N = np.array(c.n_coarse[:c.dim]) * 2 ** c.refine
    lo, hi = np.array(c.lo[:c.dim]), np.array(c.hi[:c.dim])
    ijk = np.floor((ctr - lo) / (hi - lo) * N).astype(int)
    return ijk.sum(axis=1) % 2


CASES = {
    "3d_k3_pert_cellular": dict(dim=3, degree=3, refine=2, perturbation=0.1,
                                velocity_kind=1, factor_skew=0.5),
    "3d_k4_cart": dict(dim=3, degree=4, refine=1, velocity=[1.0, 0.5, -0.25],
                       factor_skew=0.5),
    "2d_k4_vortex_pert": dict(dim=2, degree=4, refine=3, perturbation=0.1,
                              velocity_kind=3, time_factor=1, final_time=8.0,
                              factor_skew=0.5),
}
SEPARABLE = {
    "3d_k4_cart": CASES["3d_k4_cart"],
    "2d_k5_cart": dict(dim=2, degree=5, refine=2, velocity=[-0.7, 1.1],
                       factor_skew=0.5),
}


@pytest.mark.parametrize("s", [2, 3])
@pytest.mark.parametrize("name", sorted(CASES))
def test_cell_operator_matches_oracle_on_a_checkerboard(gpu_lib, name, s):
    """For values supported on cells of one parity, the global stage
    operator restricted to those cells is the cell-local operator (the
    neighbours' traces are zero); Gauss values y map to nodal values
    S^-1 y, and the weak-form result v to Gauss values S^-T v."""
    c = Cfg(irk_stages=s, **CASES[name])
    op, o = AdvectionOperation(c.dg()), c.oracle_handle()
    n, dim = c.degree + 1, c.dim
    nd = n ** dim
    S = basis_change(c.degree)
    Si, SiT = np.linalg.inv(S), np.linalg.inv(S).T
    par = cell_parity(o, c)
    t, dt = 0.2, 0.05
    op.set_time(t, dt)
    rng = np.random.default_rng(31)
    for side in (0, 1):
        y = rng.standard_normal((s, o.n_cells, nd))
        y[:, par != side, :] = 0
        u = apply3(Si, y, n, dim)
        v = o.irk_vmult(u.reshape(-1), t, dt).reshape(s, o.n_cells, nd)
        want = apply3(SiT, v, n, dim)[:, par == side, :]
        z = op.initialize_dof_vector(blocks=s)
        op.cell_operator_irk(z, dev(y.reshape(-1)))
        got = host(z).reshape(s, o.n_cells, nd)[:, par == side, :]
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err < 1e-12, (side, err)


@pytest.mark.parametrize("name", sorted(SEPARABLE))
def test_cell_solver_is_exact_for_a_separable_cell_operator(gpu_lib, name):
    """Axis-aligned cells and a constant speed: the cell operator is the
    separable one the collocation FDM inverts, so the GMRES step has
    alpha = 1 and the preconditioner solves every cell exactly. In the
    collocation basis A_c = S^-T A S^-1: the correction x satisfies
    A_c (S x) = S^-T r, and equals the nodal FDM's exact cell inverse."""
    s = 3
    c = Cfg(irk_stages=s, **SEPARABLE[name])
    op = AdvectionOperation(c.dg())
    n, dim = c.degree + 1, c.dim
    S = basis_change(c.degree)
    SiT = np.linalg.inv(S).T
    op.set_time(0.1, 0.03)
    nc = op.n_dofs // n ** dim
    r = np.random.default_rng(5).standard_normal(s * op.n_dofs)
    x = op.initialize_dof_vector(blocks=s)
    op.precondition_cell_solver_irk(x, dev(r))
    xq = apply3(S, host(x).reshape(s, nc, n ** dim), n, dim)
    rc = apply3(SiT, r.reshape(s, nc, n ** dim), n, dim)
    z = op.initialize_dof_vector(blocks=s)
    op.cell_operator_irk(z, dev(xq.reshape(-1)))
    # the bound of the nodal FDM's exactness test: the stage and 1D
    # eigenvector transforms have condition numbers of 1e2..1e4
    err = np.abs(host(z) - rc.reshape(-1)).max() / np.abs(rc).max()
    assert err < 1e-10, err
    xf = op.initialize_dof_vector(blocks=s)
    op.precondition_block_jacobi_irk(xf, dev(r))
    err = ((x - xf).abs().max() / xf.abs().max()).item()
    assert err < 1e-10, err


def test_cell_solver_first_call_on_a_fresh_operator(gpu_lib):
    """The first call sets up the stage coefficients itself (the averaged
    speed factor depends on the stage count): a zero residual gives a zero
    correction, not the 0/0 of an empty stage table."""
    c = Cfg(irk_stages=2, **CASES["3d_k4_cart"])
    op = AdvectionOperation(c.dg())
    op.set_time(0.1, 0.03)
    x = op.initialize_dof_vector(blocks=2)
    x.fill_(7.0)
    op.precondition_cell_solver_irk(x, torch.zeros_like(x))
    assert torch.count_nonzero(x).item() == 0


@pytest.mark.parametrize("name", ["3d_k3_pert_cellular", "2d_k4_vortex_pert"])
def test_cell_solver_step_minimizes_each_cell_residual(gpu_lib, name):
    """x = alpha P^-1 r with alpha minimizing |r_c - A_c x| per cell in the
    collocation basis (r_c = S^-T r, A_c = S^-T A S^-1): the cell residual
    r_c - A_c S x is orthogonal to A_c S x on every cell."""
    s = 3
    c = Cfg(irk_stages=s, **CASES[name])
    op = AdvectionOperation(c.dg())
    n, dim = c.degree + 1, c.dim
    nd = n ** dim
    S = basis_change(c.degree)
    SiT = np.linalg.inv(S).T
    op.set_time(0.3, 0.04)
    nc = op.n_dofs // nd
    r = np.random.default_rng(7).standard_normal(s * op.n_dofs)
    x = op.initialize_dof_vector(blocks=s)
    op.precondition_cell_solver_irk(x, dev(r))
    xq = apply3(S, host(x).reshape(s, nc, nd), n, dim)
    rq = apply3(SiT, r.reshape(s, nc, nd), n, dim)
    z = op.initialize_dof_vector(blocks=s)
    op.cell_operator_irk(z, dev(xq.reshape(-1)))
    Ax = host(z).reshape(s, nc, nd)
    e = rq - Ax
    dots = np.einsum("scq,scq->c", e, Ax)
    scale = np.sqrt(np.einsum("scq,scq->c", e, e) *
                    np.einsum("scq,scq->c", Ax, Ax))
    assert np.all(np.abs(dots) <= 1e-10 * scale + 1e-300)
    # and the step reduces every cell residual
    assert np.all(np.einsum("scq,scq->c", e, e) <=
                  np.einsum("scq,scq->c", rq, rq) * (1 + 1e-12))


@pytest.mark.parametrize("kw", [dict(prec_precision=L.DG_FP64),
                                dict(prec_precision=L.DG_FP32,
                                     inner_iterations=4)])
def test_irk_solve_with_the_cell_solver(gpu_lib, kw):
    s = 3
    c = Cfg(dim=3, degree=4, refine=2, perturbation=0.1, velocity_kind=1,
            factor_skew=0.5, irk_stages=s, solution_kind=3,
            solution_param=[0.5, 0.5, 0.5, 0.1])
    op = AdvectionOperation(c.dg())
    u = op.initialize_dof_vector()
    op.project_initial(u, 0.0)
    op.set_time(0.0, 0.2)
    rhs = op.initialize_dof_vector(blocks=s)
    op.compute_rhs(rhs, u)
    sols, its = [], []
    for cell_solver in (False, True):
        z = op.initialize_dof_vector(blocks=s)
        ctl = solver_control(max_iterations=200, rel_tolerance=1e-10,
                             restart=30, cell_solver=cell_solver, **kw)
        op.irk_solve(z, rhs, ctl)
        res = op.initialize_dof_vector(blocks=s)
        op.vmult_irk(res, z)
        res -= rhs
        assert (res.norm() / rhs.norm()).item() < 1e-10
        sols.append(z)
        its.append(ctl.iterations)
    diff = ((sols[0] - sols[1]).norm() / sols[0].norm()).item()
    assert diff < 1e-8, diff
    # the per-cell GMRES step is at least as good a cell inverse as the
    # FDM it is preconditioned with (102 vs 96 and 26 vs 26 iterations
    # when measured): no more than a few extra outer iterations
    assert its[1] <= its[0] + 5, its
