"""Pins the CPU oracle against closed forms, invariants and brute force.

The reference ships no tests or golden vectors (it has no test suite), so
every check here comes from mathematics that fixes the answer."""
import itertools

import numpy as np
import pytest

import oracle
from helpers import (Cfg, gl_nodes, nodal_points, rel_diff,
                     transport_error, transport_problem)


@pytest.mark.parametrize("n", range(2, 9))
def test_basis_1d_against_numpy(n):
    gp, gw, gl = oracle.basis_1d(n)
    x, w = np.polynomial.legendre.leggauss(n)
    assert np.allclose(gp, (x + 1) / 2, atol=1e-15, rtol=0)
    assert np.allclose(gw, w / 2, atol=1e-15, rtol=0)
    assert np.allclose(gl, gl_nodes(n), atol=2e-15, rtol=0)


def _refine_order(dim, n_coarse, r):
    """Cell coordinates in the order deal.II's refine_global produces on
    lexicographic coarse cells: every cell is replaced by its children
    (child c at offset (c&1, c>>1&1, c>>2&1)), recursively."""
    cells = []
    for k in range(n_coarse[2] if dim == 3 else 1):
        for j in range(n_coarse[1]):
            for i in range(n_coarse[0]):
                cells.append((i, j, k))
    for _ in range(r):
        new = []
        for (i, j, k) in cells:
            for c in range(2 ** dim):
                new.append((2 * i + (c & 1), 2 * j + ((c >> 1) & 1),
                            2 * k + ((c >> 2) & 1) if dim == 3 else 0))
        cells = new
    return np.array(cells, dtype=np.int32)


@pytest.mark.parametrize("dim,nc,r", [(2, [1, 1, 1], 3), (3, [1, 1, 1], 2),
                                      (3, [3, 1, 2], 1), (2, [3, 3, 1], 2)])
def test_cell_order_is_refine_global(dim, nc, r):
    o = Cfg(dim=dim, degree=1, n_coarse=nc, refine=r).oracle_handle()
    assert np.array_equal(o.cell_coordinates(), _refine_order(dim, nc, r))


def test_neighbors_and_periodicity():
    c = Cfg(dim=3, degree=1, n_coarse=[2, 1, 1], refine=1,
            periodic=[1, 0, 1])
    o = c.oracle_handle()
    ijk = o.cell_coordinates()
    nb = o.neighbors()
    N = [4, 2, 2]
    index = {tuple(v): i for i, v in enumerate(ijk)}
    for cell, v in enumerate(ijk):
        for d in range(3):
            for s in range(2):
                w = list(v)
                w[d] += 1 if s else -1
                if w[d] < 0 or w[d] >= N[d]:
                    if not c.periodic[d]:
                        assert nb[cell, 2 * d + s] == -1
                        continue
                    w[d] %= N[d]
                assert nb[cell, 2 * d + s] == index[tuple(w)]


def test_perturbed_vertices_seeded_and_bounded():
    c = Cfg(dim=3, degree=1, refine=3, perturbation=0.1, seed=7)
    o = c.oracle_handle()
    v = o.cell_vertices()
    h = 1.0 / 8
    ijk = o.cell_coordinates()
    for cell in range(0, len(ijk), 37):
        for corner in range(8):
            ref = (ijk[cell] + [(corner >> d) & 1 for d in range(3)]) * h
            off = v[cell, corner] - ref
            assert np.all(np.abs(off) <= 0.1 * h + 1e-15)
            on_boundary = np.any((ref == 0) | (np.abs(ref - 1) < 1e-15))
            if on_boundary:
                assert np.all(off == 0)
    # the same seed gives the same mesh, another seed another one
    v2 = Cfg(dim=3, degree=1, refine=3, perturbation=0.1,
             seed=7).oracle_handle().cell_vertices()
    v3 = Cfg(dim=3, degree=1, refine=3, perturbation=0.1,
             seed=8).oracle_handle().cell_vertices()
    assert np.array_equal(v, v2) and not np.array_equal(v, v3)


def _poly_problem(dim, k, pert):
    a = [0.7, -0.4, 0.3][:dim] + [0.0] * (3 - dim)
    return Cfg(dim=dim, degree=k, refine=1 if dim == 3 else 2,
               periodic=[0, 0, 0], perturbation=pert, velocity=a,
               solution_kind=4, solution_param=[k], lo=[-0.5, 0, 0.25],
               hi=[0.5, 1.5, 1.0])


@pytest.mark.parametrize("dim,k", [(2, 1), (2, 3), (3, 2), (3, 4)])
def test_known_answer_polynomial(dim, k):
    """u in Q_k, continuous, Dirichlet data = u: M^-1(-L u) = -a.grad u
    exactly on affine cells (QGauss(k+1) integrates degree 2k+1)."""
    c = _poly_problem(dim, k, 0.0)
    o = c.oracle_handle()
    u = o.project(0.0)
    du = o.apply(u, 0.0)
    pts = nodal_points(o.cell_vertices(), k + 1, dim)
    a = np.array(c.velocity[:dim])
    s = 1 + sum((d + 1) * pts[..., d] for d in range(dim))
    grad_dot = k * s ** (k - 1) * sum((d + 1) * a[d] for d in range(dim))
    assert rel_diff(du, -grad_dot.ravel()) < 1e-11
    # the projection itself interpolates the polynomial exactly
    assert rel_diff(u, (s ** k).ravel()) < 1e-12


@pytest.mark.parametrize("dim,k", [(2, 1), (2, 3), (3, 2), (3, 4)])
def test_gl_quadrature_known_answer_polynomial(dim, k):
    """use_gl_quad (advection_solver.cc:110-112): with QGaussLobatto(k+1)
    the quadrature points are the nodes. On axis-aligned cells every term
    of the x part reduces to w_b w_c times a line integral along x whose
    integrand (phi' u or phi u', degree 2k-1) the rule integrates exactly,
    and the lumped mass matrix is diag(w_a w_b w_c |K|): M^-1(-L u) =
    -a.grad u at the nodes for u in Q_k (summation by parts), although the
    mass matrix itself is inexact."""
    c = _poly_problem(dim, k, 0.0)
    c.use_gl_quad = 1
    o = c.oracle_handle()
    u = o.project(0.0)
    du = o.apply(u, 0.0)
    pts = nodal_points(o.cell_vertices(), k + 1, dim)
    a = np.array(c.velocity[:dim])
    s = 1 + sum((d + 1) * pts[..., d] for d in range(dim))
    grad_dot = k * s ** (k - 1) * sum((d + 1) * a[d] for d in range(dim))
    assert rel_diff(du, -grad_dot.ravel()) < 1e-11
    assert rel_diff(u, (s ** k).ravel()) < 1e-12
    # the Gauss-Lobatto mass matrix is not the exact one (degree 2k)
    g = _poly_problem(dim, k, 0.0).oracle_handle()
    assert rel_diff(o.mass(u), g.mass(u)) > 1e-6


def _gl_weights(n):
    """Gauss-Lobatto weights on [0,1] in closed form, 1 / (n (n-1)
    P_{n-1}(x_i)^2), at the roots of P'_{n-1} and the end points."""
    from numpy.polynomial import legendre as leg
    c = np.zeros(n)
    c[-1] = 1
    x = np.concatenate([[-1.0], np.sort(leg.legroots(leg.legder(c))), [1.0]])
    w = 1.0 / (n * (n - 1) * leg.legval(x, c) ** 2)
    return (x + 1) / 2, w


@pytest.mark.parametrize("dim,k,pert", [(2, 3, 0.0), (3, 4, 0.0),
                                        (3, 3, 0.15)])
def test_gl_mass_inverse_closed_form(dim, k, pert):
    """use_gl_quad_mass (advection_solver.cc:114-116): the mass matrix of
    QGaussLobatto(k+1) is diagonal in the nodal basis, so its inverse is
    v / (w_a w_b (w_c) det J) with det J at the nodes, from the closed-form
    weights and the multilinear map of the exported corners."""
    n = k + 1
    c = Cfg(dim=dim, degree=k, refine=1, perturbation=pert, seed=5,
            use_gl_quad_mass=1, lo=[-0.5, 0, 0.25], hi=[0.5, 1.5, 1.0])
    o = c.oracle_handle()
    v = np.random.default_rng(2).standard_normal(o.n_dofs)
    x, w = _gl_weights(n)
    verts = o.cell_vertices().reshape(o.n_cells, 2 ** dim, dim)
    jxw = np.empty((o.n_cells, n ** dim))
    for p in range(n ** dim):
        idx = [(p // n ** d) % n for d in range(dim)]
        xi = [x[i] for i in idx]
        # multilinear map: J[e][d] = d x_e / d xi_d
        J = np.zeros((o.n_cells, dim, dim))
        for vtx in range(2 ** dim):
            bits = [(vtx >> d) & 1 for d in range(dim)]
            for d in range(dim):
                dphi = 1.0
                for e in range(dim):
                    if e == d:
                        dphi *= 1.0 if bits[e] else -1.0
                    else:
                        dphi *= xi[e] if bits[e] else 1 - xi[e]
                J[:, :, d] += verts[:, vtx, :] * dphi
        jxw[:, p] = np.linalg.det(J) * np.prod([w[i] for i in idx])
    assert rel_diff(o.inverse_mass(v), (v.reshape(jxw.shape) / jxw).ravel()) \
        < 1e-13
    assert rel_diff(o.mass(o.inverse_mass(v)), v) < 1e-13


def test_gl_mass_apply_is_gauss_apply_remassed():
    """With use_gl_quad_mass the operator keeps QGauss(k+1) and only the
    inverse mass changes: apply_GL(u) = M_GL^-1 M_Gauss apply_Gauss(u),
    M_Gauss from the plain problem, M_GL^-1 checked in closed form above."""
    kw = dict(dim=3, degree=3, refine=1, perturbation=0.1, velocity_kind=1)
    g = Cfg(**kw).oracle_handle()
    m = Cfg(use_gl_quad_mass=1, **kw).oracle_handle()
    u = np.random.default_rng(4).standard_normal(g.n_dofs)
    ref = m.inverse_mass(g.mass(g.apply(u, 0.1)))
    assert rel_diff(m.apply(u, 0.1), ref) < 1e-12
    assert rel_diff(m.apply(u, 0.1), g.apply(u, 0.1)) > 1e-4


def _linear_problem(dim, k, pert):
    """Globally linear u = 1 + x + 2y + 3z on perturbed (non-affine) cells
    of a box, constant speed, Dirichlet data = u."""
    return Cfg(dim=dim, degree=k, refine=2, periodic=[0, 0, 0],
               perturbation=pert, seed=11,
               velocity=[0.7, -0.4, 0.3][:dim] + [0.0] * (3 - dim),
               solution_kind=4, solution_param=[1], lo=[-0.5, 0, 0.25],
               hi=[0.5, 1.5, 1.0])


@pytest.mark.parametrize("dim,k", [(2, 1), (2, 3), (3, 2), (3, 3), (3, 4)])
def test_known_answer_linear_on_perturbed_cells(dim, k):
    """Non-affine cells: a globally linear u pulls back to Q_1 under the
    multilinear map, so it lies in the DG space and is continuous. The
    cell integrand grad_ref(phi) . adj(J) a u has degree <= k + 3 per
    reference direction (adj(J) is quadratic per direction in 3D, linear
    in 2D), and so does the face integrand: QGauss(k+1) is exact for
    k >= 2 (3D), k >= 1 (2D). Then -L u = -(a.grad u) M 1 with the same
    quadrature mass matrix, so M^-1(-L u) = -a.grad u, a constant. A
    transposed Jacobian contraction (J^-T a instead of J^-1 a) breaks it."""
    c = _linear_problem(dim, k, 0.15)
    o = c.oracle_handle()
    u = o.project(0.0)
    pts = nodal_points(o.cell_vertices(), k + 1, dim)
    s = 1 + sum((d + 1) * pts[..., d] for d in range(dim))
    assert rel_diff(u, s.ravel()) < 1e-13
    a = np.array(c.velocity[:dim])
    grad_dot = sum((d + 1) * a[d] for d in range(dim))
    du = o.apply(u, 0.0)
    # cell terms of size |u||a|/h cancel to the constant: round-off ~1e-12
    assert rel_diff(du, np.full(o.n_dofs, -grad_dot)) < 1e-10


@pytest.mark.parametrize("velocity_kind", [0, 1])
@pytest.mark.parametrize("k", [2, 3])
def test_perturbed_mesh_transport_converges(k, velocity_kind):
    """3D cells perturbed by 10% of h (seeded, a new perturbation on every
    mesh): M^-1(-L u) of u = prod sin(2 pi x_d) approaches -a.grad u at
    about order k over refine 2 -> 4, for a constant speed and for the
    divergence-free cellular field (where -div(a u) = -a.grad u). With J^-T
    in place of J^-1 in the cell term the error grows instead (measured:
    0.65 -> 1.3 -> 2.4 at k = 2)."""
    errs = []
    for r in (2, 3, 4):
        o = transport_problem(k, r, velocity_kind).oracle_handle()
        du = o.apply(o.project(0.0), 0.0)
        errs.append(transport_error(du, o.cell_vertices(), k, velocity_kind))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > k - 0.75), (errs, rates)
    assert errs[-1] < 0.02 if k == 2 else errs[-1] < 2e-3, errs


def test_inverse_mass_is_inverse_of_mass():
    c = Cfg(dim=3, degree=3, refine=1, perturbation=0.15)
    o = c.oracle_handle()
    r = np.random.default_rng(1).standard_normal(o.n_dofs)
    assert rel_diff(o.mass(o.inverse_mass(r)), r) < 1e-12


def test_mass_matrix_brute_force():
    """Cell mass matrix by direct quadrature with Lagrange polynomials on
    the (numpy) Gauss-Lobatto nodes, 2D affine cell."""
    k = 3
    c = Cfg(dim=2, degree=k, refine=0, lo=[0, 0], hi=[2.0, 0.5],
            periodic=[0, 0, 0])
    o = c.oracle_handle()
    n = k + 1
    xg = gl_nodes(n)
    xq, wq = np.polynomial.legendre.leggauss(n)
    xq, wq = (xq + 1) / 2, wq / 2

    def phi(i, x):
        return np.prod([(x - xg[j]) / (xg[i] - xg[j]) for j in range(n)
                        if j != i], axis=0)
    P = np.array([[phi(i, x) for i in range(n)] for x in xq])  # [q][i]
    M1 = P.T @ np.diag(wq) @ P
    M = np.kron(M1 * 0.5, M1 * 2.0)  # y-extent 0.5 (slow), x-extent 2 (fast)
    u = np.random.default_rng(3).standard_normal(n * n)
    assert rel_diff(o.mass(u), M @ u) < 1e-13


@pytest.mark.parametrize("alpha", [0.0, 1.0])
def test_conservation(alpha):
    """Periodic mesh, conservative form: 1^T M du/dt = 0."""
    c = Cfg(dim=3, degree=3, refine=1, perturbation=0.1, velocity_kind=1,
            flux_alpha=alpha)
    o = c.oracle_handle()
    u = np.random.default_rng(2).standard_normal(o.n_dofs)
    du = o.loop(u, 0.0)      # = M * (du/dt)
    assert abs(du.sum()) < 1e-11 * np.abs(du).sum()


@pytest.mark.parametrize("alpha", [0.0, 1.0])
def test_energy_skew_form(alpha):
    """Skew-symmetric form (factor_skew = 1/2): the volume terms cancel at
    every quadrature point, so u^T M du/dt = -1/2 sum alpha |a.n| [u]^2:
    zero for the central flux, negative for upwind."""
    c = Cfg(dim=3, degree=3, refine=1, perturbation=0.1, velocity_kind=1,
            flux_alpha=alpha, factor_skew=0.5)
    o = c.oracle_handle()
    u = np.random.default_rng(2).standard_normal(o.n_dofs)
    du = o.loop(u, 0.0)
    e = u @ du
    if alpha == 0:
        assert abs(e) < 1e-12 * np.abs(u * du).sum()
    else:
        assert e < -1e-3 * np.abs(u * du).sum()


def test_constant_is_steady():
    c = Cfg(dim=2, degree=4, refine=2, perturbation=0.2, velocity_kind=2,
            time_factor=1, final_time=8.0)
    o = c.oracle_handle()
    u = np.ones(o.n_dofs)
    # the vortex is divergence free only in the continuous sense; for a
    # constant state the conservative DG form gives sum of flux jumps = 0
    # and the volume term int grad v . a; check the mass instead
    du = o.loop(u, 0.3)
    assert abs(du.sum()) < 1e-12 * np.abs(du).sum() + 1e-14


def test_linearity():
    c = Cfg(dim=3, degree=2, refine=1, perturbation=0.1, velocity_kind=1)
    o = c.oracle_handle()
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal((2, o.n_dofs))
    lhs = o.apply(2.0 * u - 3.0 * v, 0.0)
    rhs = 2.0 * o.apply(u, 0.0) - 3.0 * o.apply(v, 0.0)
    assert rel_diff(lhs, rhs) < 1e-12


def test_mass_energy_of_projection():
    c = Cfg(dim=2, degree=5, refine=2, solution_kind=2,
            solution_param=[np.pi, 0.0], lo=[-1, -1], hi=[1, 1])
    o = c.oracle_handle()
    u = o.project(0.0)
    m, e, h1 = o.mass_energy(u)
    # int sin(pi x) sin(pi y) = 0, int sin^2 sin^2 = 1, int |grad|^2 = 2 pi^2
    assert abs(m) < 1e-12
    assert abs(e - 1.0) < 1e-6
    assert abs(h1 - 2 * np.pi ** 2) < 2e-3 * 2 * np.pi ** 2
    err, nrm = o.l2_error(u, 0.0)
    assert abs(nrm - 1.0) < 1e-6 and err < 1e-5


def test_radau_iia_closed_form():
    # s = 2: c = (1/3, 1), A = [[5/12, -1/12], [3/4, 1/4]]
    A, b, c = oracle.irk_coefficients(2)
    assert np.allclose(A, [[5 / 12, -1 / 12], [3 / 4, 1 / 4]], atol=1e-15)
    A, b, c = oracle.irk_coefficients(3)
    s6 = np.sqrt(6)
    assert np.allclose(c, [(4 - s6) / 10, (4 + s6) / 10, 1.0], atol=1e-15)
    assert np.allclose(A.sum(axis=1), c, atol=1e-15)
    # reference literals (advection_solver_irk_real.cc:365-373) are exact
    ref = np.array([[0.19681547722366044, -0.06553542585019838,
                     0.02377097434822015],
                    [0.3944243147390873, 0.29207341166522843,
                     -0.04154875212599792],
                    [0.37640306270046725, 0.5124858261884216, 1. / 9.]])
    assert np.max(np.abs(A - ref)) < 2e-16
    # eigenvalues of A^-1 (Radau IIA(3))
    ev = np.sort_complex(np.linalg.eigvals(np.linalg.inv(A)))
    assert np.allclose(ev, np.sort_complex(np.array(
        [3.6378342527, 2.6810828736 + 3.0504301992j,
         2.6810828736 - 3.0504301992j])), atol=1e-9)
    # order conditions B(2s-1) for s = 4: sum b c^(k-1) = 1/k
    A, b, c = oracle.irk_coefficients(4)
    for kk in range(1, 8):
        assert abs(b @ c ** (kk - 1) - 1 / kk) < 1e-13


def test_lsrk_order_of_accuracy():
    """LSRK(5,4): the difference to a run with a 16x smaller time step (same
    spatial discretization) drops by ~2^4 per halving of the time step."""
    c = Cfg(dim=2, degree=3, refine=2, lo=[-1, -1], hi=[1, 1],
            velocity=[1.0, 0.15], solution_kind=2, solution_param=[np.pi, 0])
    o = c.oracle_handle()
    T = 0.2

    def run(nsteps):
        dt = T / nsteps
        u = o.project(0.0)
        r1, r2 = o.zeros(), o.zeros()
        t = 0.0
        for _ in range(nsteps):
            o.lsrk_time_step(1, t, dt, u, r1, r2)
            t += dt
        return u
    ref = run(256)
    e1 = np.linalg.norm(run(8) - ref)
    e2 = np.linalg.norm(run(16) - ref)
    rate = np.log2(e1 / e2)
    assert 3.7 < rate < 4.5, (e1, e2, rate)


def test_sipg_symmetric_and_coercive():
    """Pure diffusion (a = 0), positive operator: SIPG is symmetric and,
    with the reference's penalty 2k(k+1), coercive."""
    k = 3
    c = Cfg(dim=3, degree=k, refine=1, velocity=[0, 0, 0], operator_kind=1,
            viscosity=0.7, penalty_factor=2 * k * (k + 1))
    o = c.oracle_handle()
    rng = np.random.default_rng(9)
    u, v = rng.standard_normal((2, o.n_dofs))
    Lu, Lv = o.vmult(u, 0.0, 0.0), o.vmult(v, 0.0, 0.0)
    assert abs(v @ Lu - u @ Lv) < 1e-12 * abs(v @ Lu)
    assert u @ Lu > 0 and v @ Lv > 0


def _sipg_perturbed(k, refine, pert, **kw):
    return Cfg(dim=3, degree=k, refine=refine, perturbation=pert, seed=7,
               velocity=[0, 0, 0], operator_kind=1,
               penalty_factor=2 * k * (k + 1), **kw)


def test_sipg_symmetric_and_coercive_on_perturbed_cells():
    """Pure diffusion on cells perturbed by 15% of h: the SIPG form stays
    symmetric and coercive. Contracting the reference gradients with J^-1
    instead of J^-T breaks the symmetry (relative 3e-2 on a scratch copy of
    the oracle, against 1e-15)."""
    k = 3
    o = _sipg_perturbed(k, 1, 0.15, viscosity=0.7).oracle_handle()
    rng = np.random.default_rng(9)
    u, v = rng.standard_normal((2, o.n_dofs))
    Lu, Lv = o.vmult(u, 0.0, 0.0), o.vmult(v, 0.0, 0.0)
    assert abs(v @ Lu - u @ Lv) < 1e-12 * abs(v @ Lu)
    assert u @ Lu > 0 and v @ Lv > 0


@pytest.mark.parametrize("k", [3, 4])
def test_sipg_perturbed_mesh_converges(k):
    """Pure diffusion (nu = 1) on 3D cells perturbed by 10% of h: M^-1(-L u)
    of u = prod sin(2 pi x_d) approaches nu lap u = -3 (2 pi)^2 u at about
    order k - 1 (RMS over the nodes, refine 1 -> 3). The scratch oracle with
    J^-1 in place of J^-T diverges (rates -0.04, -1.2 at k = 3)."""
    errs = []
    for r in (1, 2, 3):
        o = _sipg_perturbed(k, r, 0.1, viscosity=1.0, solution_kind=2,
                            solution_param=[2 * np.pi, 0.0]).oracle_handle()
        du = o.apply(o.project(0.0), 0.0)
        p = nodal_points(o.cell_vertices(), k + 1, 3)
        ex = -3 * (2 * np.pi) ** 2 * np.prod(np.sin(2 * np.pi * p), -1)
        ex = ex.ravel()
        errs.append(np.sqrt(np.mean((du - ex) ** 2) / np.mean(ex ** 2)))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all(rates > k - 2), (errs, rates)
    assert errs[-1] < (0.03 if k == 3 else 3e-3), errs


def test_sipg_closed_form_decay():
    """Explicit advection-diffusion right-hand side of the product of sines
    approaches -a.grad u + nu lap u (spectral accuracy in k)."""
    nu, k = 0.05, 6
    a = [1.0, 0.5, 0.25]
    c = Cfg(dim=3, degree=k, refine=1, velocity=a, operator_kind=1,
            viscosity=nu, penalty_factor=(k + 1) ** 2, solution_kind=2,
            solution_param=[2 * np.pi, nu])
    o = c.oracle_handle()
    u = o.project(0.0)
    du = o.apply(u, 0.0)
    # exact d/dt of prod sin(2 pi (x_d - a_d t)) exp(-3 nu (2 pi)^2 t) at t=0
    # is projected with the same interpolation as u
    h = 1e-5
    ddt = (o.project(h) - o.project(-h)) / (2 * h)
    assert rel_diff(du, ddt) < 2e-3


def _burgers(dim=3, k=3, refine=1, **kw):
    return Cfg(dim=dim, degree=k, n_components=dim, operator_kind=2,
               refine=refine, lo=[-1, -1, -1], hi=[1, 1, 1],
               solution_kind=6, **kw)


@pytest.mark.parametrize("k", [3, 6])
def test_burgers_constant_state_is_steady(k):
    o = _burgers(k=k, perturbation=0.1).oracle_handle()
    u = np.zeros(o.n_dofs * 3)
    nd = (k + 1) ** 3
    for c, val in enumerate((0.7, -0.3, 1.1)):
        u.reshape(-1, 3, nd)[:, c, :] = val
    du = o.burgers_apply(u)
    # round-off of the inverse mass grows with k (2e-13 at k = 3, 1.5e-12
    # at k = 6); an operator that does not keep the state steady gives O(1)
    assert np.max(np.abs(du)) < 1e-11


def test_burgers_smooth_right_hand_side():
    """u = (-sin(pi x), 0, 0): du/dt = -d/dx(u_0 u_0) = -pi sin(2 pi x)."""
    o = _burgers(k=4, refine=2).oracle_handle()
    u = o.project(0.0)
    du = o.burgers_apply(u).reshape(-1, 3, 125)
    pts = nodal_points(o.cell_vertices(), 5, 3)
    exact = -np.pi * np.sin(2 * np.pi * pts[..., 0])
    assert rel_diff(du[:, 0, :], exact) < 2e-2
    assert np.max(np.abs(du[:, 1:, :])) < 1e-12


def test_burgers_conserves_momentum():
    o = _burgers(perturbation=0.1).oracle_handle()
    u = np.random.default_rng(3).standard_normal(o.n_dofs * 3)
    du = o.burgers_apply(u).reshape(-1, 3, 64)
    for c in range(3):
        comp = np.ascontiguousarray(du[:, c, :]).ravel()
        m = o.mass(comp)
        assert abs(m.sum()) < 1e-11 * np.abs(m).sum()


# LSRK(5,4) of Kennedy, Carpenter and Lewis (2000), the rationals of
# advection_solver.cc:840-848
LSRK54_B = [1153189308089. / 22510343858157., 1772645290293. / 4653164025191.,
            -1672844663538. / 4480602732383., 2114624349019. / 3568978502595.,
            5198255086312. / 14908931495163.]
LSRK54_A = [970286171893. / 4311952581923., 6584761158862. / 12103376702013.,
            2251764453980. / 15575788980749., 26877169314380. / 34165994151039.]


def _burgers_exact(x, t):
    """u = -sin(pi (x - 2 u t)), the solution of u_t + (u u)_x = 0 (the
    first component of the flux u (x) u, burgers_solver.cc:277-280) from
    u0 = -sin(pi x) along the characteristics of speed 2u, by Newton per
    point (t below the break time 1/(2 pi))."""
    u = -np.sin(np.pi * x)
    for _ in range(60):
        a = np.pi * (x - 2.0 * u * t)
        u = u - (u + np.sin(a)) / (1.0 - 2.0 * np.pi * t * np.cos(a))
    return u


@pytest.mark.parametrize("k", [2, 3])
def test_burgers_converges_to_characteristic_solution(k):
    """2D Burgers from u0 = (-sin(pi x), 0) on [-1,1]^2, periodic, to
    t = 0.05 (a third of the way to the shock at 1/(2 pi)): the L2 error
    against the exact solution by characteristics falls at order k+1 under
    refinement (measured: k=2 2.7, 3.1; k=3 4.1, 3.9), and the second
    component stays zero. The time loop restates the two-register recursion
    of advection_solver.cc:906-939 over the oracle's right-hand side."""
    T, errs = 0.05, []
    for r in (3, 4, 5):
        o = _burgers(dim=2, k=k, refine=r).oracle_handle()
        u = o.project(0.0)
        h = 2.0 / 2 ** r
        n = int(np.ceil(T / (0.25 * h / (k + 1) ** 1.5)))
        dt = T / n
        for _ in range(n):
            ri = u
            for i, b in enumerate(LSRK54_B):
                kv = o.burgers_apply(ri)
                if i < len(LSRK54_A):
                    ri = u + LSRK54_A[i] * dt * kv
                u = u + b * dt * kv
        nd = (k + 1) ** 2
        uc = u.reshape(-1, 2, nd)
        pts = nodal_points(o.cell_vertices(), k + 1, 2)
        # L2 error of the first component against the exact solution's
        # nodal interpolant, with the oracle's (pinned) mass matrix
        e = np.ascontiguousarray(uc[:, 0, :] - _burgers_exact(pts[..., 0],
                                                              T)).ravel()
        errs.append(np.sqrt(e @ o.mass(e)))
        assert np.max(np.abs(uc[:, 1, :])) < 1e-13
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert errs[-1] < 2e-4 and min(rates) > k + 0.5, (errs, rates)


# ---- implicit and IRK operators: compositions of pinned pieces ------------

IMPLICIT = dict(dim=2, degree=3, refine=2, perturbation=0.15,
                velocity_kind=2, time_factor=1, final_time=4.0,
                factor_skew=0.5)


def test_implicit_vmult_is_mass_over_dt_plus_operator():
    """(M/dt + L(t+dt)) u from its two independently pinned parts, the mass
    matrix and the explicit loop (which returns -L u)."""
    o = Cfg(**IMPLICIT).oracle_handle()
    u = np.random.default_rng(11).standard_normal(o.n_dofs)
    t, dt = 0.3, 0.07
    ref = o.mass(u) / dt - o.loop(u, t + dt)
    assert rel_diff(o.vmult(u, t, dt), ref) < 1e-13


@pytest.mark.parametrize("s", [2, 3])
def test_irk_vmult_is_stage_coupled_mass_plus_operator(s):
    """Block c of the IRK operator: sum_b (A^-1)_cb M Z_b / dt + L(t+c_c dt)
    Z_c, assembled from the mass matrix, the explicit loop and the Radau IIA
    coefficients."""
    o = Cfg(irk_stages=s, **IMPLICIT).oracle_handle()
    N = o.n_dofs
    Z = np.random.default_rng(12).standard_normal(N * s).reshape(s, N)
    t, dt = 0.2, 0.05
    A, b, c = oracle.irk_coefficients(s)
    Ainv = np.linalg.inv(A)
    MZ = np.array([o.mass(z) for z in Z])
    ref = np.array([Ainv[i] @ MZ / dt - o.loop(Z[i], t + c[i] * dt)
                    for i in range(s)])
    assert rel_diff(o.irk_vmult(Z.ravel(), t, dt), ref.ravel()) < 1e-13
    rhs = np.array([o.loop(Z[0], t + c[i] * dt) / dt for i in range(s)])
    assert rel_diff(o.irk_compute_rhs(Z[0], t, dt), rhs.ravel()) < 1e-13


def test_irk_step_converges_with_order_2s_minus_1():
    """Radau IIA(3) steps of the oracle's compute_rhs / vmult /
    update_solution against the exact solution exp(-T M^-1 L) u0 of the
    semi-discrete system: the error falls like dt^5."""
    from scipy.linalg import expm
    s = 3
    c = Cfg(dim=2, degree=2, refine=1, velocity=[1.0, 0.5], factor_skew=0.5,
            irk_stages=s, solution_kind=2, solution_param=[2 * np.pi, 0])
    o = c.oracle_handle()
    N = o.n_dofs
    u0 = o.project(0.0)
    Mh = np.column_stack([o.apply(e, 0.0) for e in np.eye(N)])
    T = 0.4
    exact = expm(T * Mh) @ u0
    errs = []
    for nsteps in (4, 8, 16):
        dt = T / nsteps
        Big = np.column_stack([o.irk_vmult(e, 0.0, dt)
                               for e in np.eye(N * s)])
        u = u0.copy()
        for _ in range(nsteps):
            Z = np.linalg.solve(Big, o.irk_compute_rhs(u, 0.0, dt))
            o.irk_update_solution(u, Z, dt)
        errs.append(np.linalg.norm(u - exact))
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert np.all((rates > 4.5) & (rates < 5.6)), (errs, rates)
