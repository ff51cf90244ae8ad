"""VTU output (advection_solver.cc:1217-1241): node order of the VTK
Lagrange cells, point positions, the solution evaluated at the patch points
and the interpolated analytic solution; host writer on the CPU, device
entry point against it on the GPU."""
import ctypes
import re

import numpy as np
import pytest

from advection_miniapp_amd import _lib
from helpers import Cfg


def read_vtu(path):
    raw = open(path, "rb").read()
    start = raw.index(b'<AppendedData encoding="raw">')
    head = raw[:start].decode()
    base = raw.index(b"_", start) + 1
    arrays = {}
    for m in re.finditer(r'<DataArray type="(\w+)"(?: Name="(\w+)")? '
                         r'NumberOfComponents="(\d+)" format="appended" '
                         r'offset="(\d+)"/>', head):
        typ, name, comps, off = m.group(1), m.group(2) or "Points", \
            int(m.group(3)), int(m.group(4))
        nbytes = int(np.frombuffer(raw, np.uint64, 1, base + off)[0])
        dt = {"Float64": np.float64, "Int64": np.int64, "UInt8": np.uint8}[typ]
        a = np.frombuffer(raw, dt, nbytes // np.dtype(dt).itemsize,
                          base + off + 8)
        arrays[name] = a.reshape(-1, comps) if comps > 1 else a
    npts = int(re.search(r'NumberOfPoints="(\d+)"', head).group(1))
    ncel = int(re.search(r'NumberOfCells="(\d+)"', head).group(1))
    assert 'version="2.2"' in head
    return arrays, npts, ncel


def gl_nodes(k):
    """Gauss-Lobatto nodes on [0, 1] (roots of P_k' and the end points)."""
    from numpy.polynomial import legendre
    inner = legendre.Legendre.basis(k).deriv().roots() if k > 1 else []
    return (np.concatenate(([-1.0], np.sort(inner), [1.0])) + 1) / 2


def vtk_order(dim, q):
    """Lattice points (i, j, l) of a VTK Lagrange quadrilateral/hexahedron
    of order q in VTK's node order (file format 2.2): corners as the linear
    cell, then the edge points in the order of the linear cell's edges
    (increasing parametric coordinate), the face points (faces -x, +x, -y,
    +y, -z, +z, each lexicographic), the interior lexicographic."""
    r = range(1, q)
    if dim == 2:
        out = [(0, 0, 0), (q, 0, 0), (q, q, 0), (0, q, 0)]
        out += [(i, 0, 0) for i in r] + [(q, j, 0) for j in r]
        out += [(i, q, 0) for i in r] + [(0, j, 0) for j in r]
        out += [(i, j, 0) for j in r for i in r]
        return out
    c = [(0, 0), (q, 0), (q, q), (0, q)]
    out = [(x, y, 0) for x, y in c] + [(x, y, q) for x, y in c]
    for z in (0, q):
        out += [(i, 0, z) for i in r] + [(q, j, z) for j in r]
        out += [(i, q, z) for i in r] + [(0, j, z) for j in r]
    for x, y in c:
        out += [(x, y, l) for l in r]
    for x in (0, q):
        out += [(x, j, l) for l in r for j in r]
    for y in (0, q):
        out += [(i, y, l) for l in r for i in r]
    for z in (0, q):
        out += [(i, j, z) for j in r for i in r]
    out += [(i, j, l) for l in r for j in r for i in r]
    return out


def write_host(c, values, t, path):
    lib = _lib.load()
    cfg = c.dg()
    vec = np.ascontiguousarray(values, dtype=np.float64)
    _lib.check(lib.dg_host_write_vtu(
        ctypes.byref(cfg), vec.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
        t, str(path).encode()))


@pytest.mark.parametrize("dim,k", [(3, 3), (3, 4), (2, 3), (3, 1)])
def test_node_order_and_points(tmp_path, dim, k):
    c = Cfg(dim=dim, degree=k, refine=1, perturbation=0.1, seed=7)
    o = c.oracle_handle()
    n = k + 1
    nd = n ** dim
    write_host(c, np.zeros(o.n_dofs), 0.0, tmp_path / "a.vtu")
    arr, npts, ncel = read_vtu(tmp_path / "a.vtu")
    assert npts == o.n_cells * nd and ncel == o.n_cells
    assert np.all(arr["types"] == (72 if dim == 3 else 70))
    assert np.array_equal(arr["offsets"], nd * np.arange(1, ncel + 1))
    conn = arr["connectivity"].reshape(ncel, nd)
    order = vtk_order(dim, k)
    lex = np.array([i + n * (j + n * l) for i, j, l in order])
    assert sorted(lex) == list(range(nd))
    for cell in range(ncel):
        assert np.array_equal(conn[cell], cell * nd + lex)
    # points: the multilinear map of the cell's corners at j / k
    V = o.cell_vertices().reshape(o.n_cells, 2 ** dim, dim)
    pts = arr["Points"].reshape(ncel, nd, 3)
    t = np.arange(n) / k
    for cell in range(ncel):
        for p in range(nd):
            xi = [t[p % n], t[p // n % n], t[p // (n * n)] if dim == 3 else 0]
            x = np.zeros(dim)
            for v in range(2 ** dim):
                w = np.prod([xi[d] if v >> d & 1 else 1 - xi[d]
                             for d in range(dim)])
                x += w * V[cell, v]
            assert np.allclose(pts[cell, p, :dim], x, rtol=0, atol=1e-14)
        if dim == 2:
            assert np.all(pts[cell, :, 2] == 0)


def test_solution_and_analytic_fields(tmp_path):
    """Nodal values of p(xi) in Q_k give p at the patch points; a linear
    analytic solution is reproduced exactly by its interpolant on the
    multilinear cells."""
    dim, k = 3, 3
    c = Cfg(dim=dim, degree=k, refine=1, perturbation=0.1, seed=3,
            solution_kind=4, solution_param=[1.0], velocity=[0.5, -0.25, 1.0])
    o = c.oracle_handle()
    n = k + 1
    nd = n ** dim
    x = gl_nodes(k)

    def p(a, b, cc, cell):
        return 1 + a * a * b - 3 * cc ** 3 + 0.5 * a * b * cc + cell

    nodal = np.array([p(x[i % n], x[i // n % n], x[i // (n * n)], cell)
                      for cell in range(o.n_cells) for i in range(nd)])
    t = 0.3
    write_host(c, nodal, t, tmp_path / "b.vtu")
    arr, npts, ncel = read_vtu(tmp_path / "b.vtu")
    e = np.arange(n) / k
    want = np.array([p(e[i % n], e[i // n % n], e[i // (n * n)], cell)
                     for cell in range(o.n_cells) for i in range(nd)])
    assert np.max(np.abs(arr["solution"] - want)) < 1e-12
    # DG_SOL_POLYNOMIAL with p0 = 1: 1 + sum_d (d+1)(x_d - a_d t)
    P = arr["Points"]
    a = np.array([0.5, -0.25, 1.0])
    exact = 1 + sum((d + 1) * (P[:, d] - a[d] * t) for d in range(dim))
    assert np.max(np.abs(arr["analytic_solution"] - exact)) < 1e-12
    assert float(arr["TIME"][0]) == t


def test_vector_valued_fields(tmp_path):
    """Burgers-type vectors (components per cell) give one field per
    component."""
    c = Cfg(dim=2, degree=2, refine=2, n_components=2, operator_kind=2)
    o = c.oracle_handle()
    nd = 9
    vals = np.concatenate([np.full(nd, 10.0 * cell + comp)
                           for cell in range(o.n_cells) for comp in range(2)])
    write_host(c, vals, 0.0, tmp_path / "c.vtu")
    arr, npts, ncel = read_vtu(tmp_path / "c.vtu")
    for comp in range(2):
        f = arr[f"solution_{comp}"].reshape(ncel, nd)
        assert np.allclose(f, 10.0 * np.arange(ncel)[:, None] + comp,
                           rtol=0, atol=1e-12)


@pytest.mark.gpu
def test_device_writer_equals_host_writer(gpu_lib, tmp_path):
    import torch
    from advection_miniapp_amd import AdvectionOperation
    c = Cfg(dim=3, degree=4, refine=2, perturbation=0.1, velocity_kind=1,
            solution_kind=3, solution_param=[0.5, 0.5, 0.5, 0.2])
    op = AdvectionOperation(c.dg())
    u = op.initialize_dof_vector()
    op.project_initial(u, 0.0)
    op.write_vtu(u, 0.0, tmp_path / "dev.vtu")
    write_host(c, u.cpu().numpy(), 0.0, tmp_path / "host.vtu")
    assert open(tmp_path / "dev.vtu", "rb").read() == \
        open(tmp_path / "host.vtu", "rb").read()
