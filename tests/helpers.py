"""Shared problem descriptions for the oracle and the GPU library."""
import ctypes

import numpy as np

DEFAULTS = dict(
    dim=3, degree=4, n_components=1, operator_kind=0, precision=0,
    n_coarse=[1, 1, 1], refine=2, lo=[0.0, 0.0, 0.0], hi=[1.0, 1.0, 1.0],
    periodic=[1, 1, 1], perturbation=0.0, seed=42, velocity_kind=0,
    velocity=[1.0, 0.0, 0.0], time_factor=0, final_time=1.0, flux_alpha=1.0,
    factor_skew=0.0, viscosity=0.0, penalty_factor=0.0, solution_kind=0,
    solution_param=[0.0] * 8, irk_stages=3, use_gl_quad=0,
    use_gl_quad_mass=0)


class Cfg:
    """Plain problem description; converts to DgConfig and OrProblem."""

    def __init__(self, **kw):
        for k, v in DEFAULTS.items():
            setattr(self, k, list(v) if isinstance(v, list) else v)
        for k, v in kw.items():
            if k not in DEFAULTS:
                raise KeyError(k)
            if isinstance(DEFAULTS[k], list):
                cur = getattr(self, k)
                for i, x in enumerate(v):
                    cur[i] = x
            else:
                setattr(self, k, v)

    def _fill(self, struct):
        for name, _ in struct._fields_:
            v = getattr(self, name)
            if isinstance(v, list):
                arr = getattr(struct, name)
                for i in range(len(arr)):
                    arr[i] = v[i] if i < len(v) else 0
            else:
                setattr(struct, name, v)
        return struct

    def dg(self):
        from advection_miniapp_amd import _lib
        return self._fill(_lib.DgConfig())

    def oracle(self):
        import oracle
        return self._fill(oracle.OrProblem())

    def oracle_handle(self):
        import oracle
        return oracle.Oracle(self)

    def __getitem__(self, k):
        return getattr(self, k)


def rel_diff(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def random_vector(n, seed=0):
    return np.random.default_rng(seed).standard_normal(n)


def gl_nodes(n):
    """Gauss-Lobatto nodes on [0,1] from numpy (roots of P'_{n-1})."""
    from numpy.polynomial import legendre as Lg
    if n == 2:
        return np.array([0.0, 1.0])
    c = np.zeros(n)
    c[-1] = 1
    r = np.sort(Lg.legroots(Lg.legder(c)).real)
    return np.concatenate([[0.0], (r + 1) / 2, [1.0]])


def nodal_points(verts, n, dim):
    """Physical coordinates of the Gauss-Lobatto nodes of every cell from
    the cell corners, lexicographic (x fastest): [cells, n^dim, dim]."""
    x = gl_nodes(n)
    if dim == 3:
        Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
        ref = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    else:
        Y, X = np.meshgrid(x, x, indexing="ij")
        ref = np.stack([X.ravel(), Y.ravel()], axis=1)
    nv = 2 ** dim
    N = np.ones((ref.shape[0], nv))
    for v in range(nv):
        for d in range(dim):
            b = (v >> d) & 1
            N[:, v] *= ref[:, d] if b else 1 - ref[:, d]
    return np.einsum("pv,cvd->cpd", N, verts)


def cellular_velocity(p):
    """(sin2πy cos2πz, sin2πz cos2πx, sin2πx cos2πy) at points [..., 3]."""
    t = 2 * np.pi
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    return np.stack([np.sin(t * y) * np.cos(t * z),
                     np.sin(t * z) * np.cos(t * x),
                     np.sin(t * x) * np.cos(t * y)], axis=-1)


def sine_product_gradient(p, w=2 * np.pi):
    """Gradient of prod_d sin(w x_d) at points [..., 3]."""
    S, C = np.sin(w * p), np.cos(w * p)
    return w * np.stack([C[..., 0] * S[..., 1] * S[..., 2],
                         S[..., 0] * C[..., 1] * S[..., 2],
                         S[..., 0] * S[..., 1] * C[..., 2]], axis=-1)


def transport_error(du, verts, k, velocity_kind, velocity=(0.7, -0.4, 0.3)):
    """Relative RMS difference at the nodes between an explicit right-hand
    side du = M^-1(-L u_h) of u = prod sin(2 pi x_d) (3D) and the exact
    -a.grad u, for the cellular (velocity_kind 1) or a constant speed."""
    p = nodal_points(verts, k + 1, 3)
    if velocity_kind == 1:
        a = cellular_velocity(p)
    else:
        a = np.broadcast_to(np.asarray(velocity, dtype=float), p.shape)
    ex = -(a * sine_product_gradient(p)).sum(-1).ravel()
    du = np.asarray(du, dtype=np.float64)
    return float(np.sqrt(np.mean((du - ex) ** 2)) / np.sqrt(np.mean(ex ** 2)))


def transport_problem(k, refine, velocity_kind, perturbation=0.1, **kw):
    """3D periodic unit cube, perturbed interior vertices, u = prod
    sin(2 pi x_d): the problem of the geometry convergence tests."""
    return Cfg(dim=3, degree=k, refine=refine, perturbation=perturbation,
               velocity_kind=velocity_kind, velocity=[0.7, -0.4, 0.3],
               solution_kind=2, solution_param=[2 * np.pi, 0.0], **kw)
