"""CPU oracle of the DG operators: test infrastructure only.

A plain C++ restatement of the reference's algorithms (oracle.cpp, built
into liboracle.so by oracle/Makefile). Only tests/, __graft_entry__.smoke()
and bench.py's CPU-baseline leg use it, as the reference of the parity
checks; the product never links or calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrProblem(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("degree", ctypes.c_int),
        ("n_components", ctypes.c_int),
        ("operator_kind", ctypes.c_int),
        ("n_coarse", ctypes.c_int * 3),
        ("refine", ctypes.c_int),
        ("lo", ctypes.c_double * 3),
        ("hi", ctypes.c_double * 3),
        ("periodic", ctypes.c_int * 3),
        ("perturbation", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("velocity_kind", ctypes.c_int),
        ("velocity", ctypes.c_double * 3),
        ("time_factor", ctypes.c_int),
        ("final_time", ctypes.c_double),
        ("flux_alpha", ctypes.c_double),
        ("factor_skew", ctypes.c_double),
        ("viscosity", ctypes.c_double),
        ("penalty_factor", ctypes.c_double),
        ("solution_kind", ctypes.c_int),
        ("solution_param", ctypes.c_double * 8),
        ("irk_stages", ctypes.c_int),
        ("use_gl_quad", ctypes.c_int),
        ("use_gl_quad_mass", ctypes.c_int),
    ]


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    sig = {
        "or_create": (_vp, [ctypes.POINTER(OrProblem)]),
        "or_destroy": (None, [_vp]),
        "or_n_cells": (ctypes.c_int64, [_vp]),
        "or_n_dofs": (ctypes.c_int64, [_vp]),
        "or_set_threads": (None, [ctypes.c_int]),
        "or_cell_coordinates": (None, [_vp, ctypes.POINTER(ctypes.c_int32)]),
        "or_neighbors": (None, [_vp, ctypes.POINTER(ctypes.c_int32)]),
        "or_cell_vertices": (None, [_vp, _dp]),
        "or_min_vertex_distance": (ctypes.c_double, [_vp]),
        "or_basis_1d": (None, [ctypes.c_int, _dp, _dp, _dp]),
        "or_apply": (None, [_vp, _dp, _dp, ctypes.c_double]),
        "or_loop": (None, [_vp, _dp, _dp, ctypes.c_double]),
        "or_inverse_mass": (None, [_vp, _dp, _dp]),
        "or_mass": (None, [_vp, _dp, _dp]),
        "or_perform_stage": (None, [_vp] + [ctypes.c_double] * 3 + [_dp] * 4),
        "or_lsrk_time_step": (None, [_vp, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, _dp, _dp, _dp]),
        "or_project": (None, [_vp, _dp, ctypes.c_double]),
        "or_mass_energy": (None, [_vp, _dp, _dp]),
        "or_l2_error": (None, [_vp, _dp, ctypes.c_double, _dp]),
        "or_vmult": (None, [_vp, ctypes.c_double, ctypes.c_double, _dp, _dp]),
        "or_irk_coefficients": (None, [ctypes.c_int, _dp, _dp, _dp]),
        "or_irk_compute_rhs": (None, [_vp, ctypes.c_double, ctypes.c_double,
                                      _dp, _dp]),
        "or_irk_vmult": (None, [_vp, ctypes.c_double, ctypes.c_double, _dp,
                                _dp]),
        "or_irk_update_solution": (None, [_vp, ctypes.c_double, _dp, _dp]),
        "or_burgers_apply": (None, [_vp, _dp, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def problem_from_config(cfg):
    """OrProblem with the same fields as a library DgConfig."""
    p = OrProblem()
    for name, _ in OrProblem._fields_:
        v = getattr(cfg, name)
        if isinstance(v, (ctypes.Array, list, tuple)):
            arr = getattr(p, name)
            for i in range(len(arr)):
                arr[i] = v[i]
        else:
            setattr(p, name, v)
    return p


def basis_1d(n):
    lib = load()
    gp, gw, gl = (np.zeros(n) for _ in range(3))
    lib.or_basis_1d(n, _p(gp), _p(gw), _p(gl))
    return gp, gw, gl


def irk_coefficients(s):
    lib = load()
    A, b, c = np.zeros(s * s), np.zeros(s), np.zeros(s)
    lib.or_irk_coefficients(s, _p(A), _p(b), _p(c))
    return A.reshape(s, s), b, c


class Oracle:
    """Operator oracle on numpy float64 vectors."""

    def __init__(self, cfg):
        self.lib = load()
        self.problem = problem_from_config(cfg)
        self.h = self.lib.or_create(ctypes.byref(self.problem))
        self.n_cells = self.lib.or_n_cells(self.h)
        self.n_dofs = self.lib.or_n_dofs(self.h)
        self.dim = cfg.dim

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_destroy(self.h)
            self.h = None

    def zeros(self, blocks=1):
        return np.zeros(self.n_dofs * blocks * self.problem.n_components)

    def apply(self, src, t):
        dst = np.zeros_like(src)
        self.lib.or_apply(self.h, _p(src), _p(dst), t)
        return dst

    def burgers_apply(self, src):
        dst = np.zeros_like(src)
        self.lib.or_burgers_apply(self.h, _p(src), _p(dst))
        return dst

    def loop(self, src, t):
        dst = np.zeros_like(src)
        self.lib.or_loop(self.h, _p(src), _p(dst), t)
        return dst

    def inverse_mass(self, src):
        dst = np.zeros_like(src)
        self.lib.or_inverse_mass(self.h, _p(src), _p(dst))
        return dst

    def mass(self, src):
        dst = np.zeros_like(src)
        self.lib.or_mass(self.h, _p(src), _p(dst))
        return dst

    def perform_stage(self, t, b_dt, a_dt, current_ri, vec_ki, solution,
                      next_ri):
        self.lib.or_perform_stage(self.h, t, b_dt, a_dt, _p(current_ri),
                                  _p(vec_ki), _p(solution), _p(next_ri))

    def lsrk_time_step(self, scheme, t, dt, solution, vec_ri, vec_ki):
        self.lib.or_lsrk_time_step(self.h, scheme, t, dt, _p(solution),
                                   _p(vec_ri), _p(vec_ki))

    def project(self, t=0.0):
        dst = self.zeros()
        self.lib.or_project(self.h, _p(dst), t)
        return dst

    def mass_energy(self, vec):
        out = np.zeros(3)
        self.lib.or_mass_energy(self.h, _p(vec), _p(out))
        return out

    def l2_error(self, vec, t):
        out = np.zeros(2)
        self.lib.or_l2_error(self.h, _p(vec), t, _p(out))
        return out

    def vmult(self, src, t, dt):
        dst = np.zeros_like(src)
        self.lib.or_vmult(self.h, t, dt, _p(src), _p(dst))
        return dst

    def irk_compute_rhs(self, src, t, dt):
        s = self.problem.irk_stages
        dst = np.zeros(src.size * s)
        self.lib.or_irk_compute_rhs(self.h, t, dt, _p(src), _p(dst))
        return dst

    def irk_vmult(self, src, t, dt):
        dst = np.zeros_like(src)
        self.lib.or_irk_vmult(self.h, t, dt, _p(src), _p(dst))
        return dst

    def irk_update_solution(self, solution, stage, dt):
        self.lib.or_irk_update_solution(self.h, dt, _p(solution), _p(stage))

    def cell_coordinates(self):
        a = np.zeros((self.n_cells, 3), dtype=np.int32)
        self.lib.or_cell_coordinates(
            self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return a

    def neighbors(self):
        a = np.zeros((self.n_cells, 2 * self.dim), dtype=np.int32)
        self.lib.or_neighbors(
            self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return a

    def cell_vertices(self):
        a = np.zeros((self.n_cells, 2 ** self.dim, self.dim))
        self.lib.or_cell_vertices(self.h, _p(a))
        return a

    def min_vertex_distance(self):
        return self.lib.or_min_vertex_distance(self.h)
