"""Shared problem descriptions for the oracle and the GPU library."""
import ctypes

import numpy as np

DEFAULTS = dict(
    dim=3, degree=4, n_components=1, operator_kind=0, precision=0,
    n_coarse=[1, 1, 1], refine=2, lo=[0.0, 0.0, 0.0], hi=[1.0, 1.0, 1.0],
    periodic=[1, 1, 1], perturbation=0.0, seed=42, velocity_kind=0,
    velocity=[1.0, 0.0, 0.0], time_factor=0, final_time=1.0, flux_alpha=1.0,
    factor_skew=0.0, viscosity=0.0, penalty_factor=0.0, solution_kind=0,
    solution_param=[0.0] * 8, irk_stages=3)


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
