// VTU output of a DG solution, as the reference's output_results writes it
// (advection_solver.cc:1217-1241): DataOut with the fields "solution" and
// "analytic_solution" (the exact solution interpolated into the DG space),
// build_patches(mapping, degree) and write_higher_order_cells. Every cell
// becomes one VTK Lagrange cell of the operator's degree k: (k+1)^d points
// at equidistant reference positions j/k, placed with the multilinear cell
// map, carrying the DG polynomial evaluated there. The file is XML
// UnstructuredGrid (format version 2.2, the Lagrange node order of current
// VTK readers) with raw appended data.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "analytic.hpp"
#include "aux.hpp"
#include "basis.hpp"
#include "internal.hpp"
#include "mesh.hpp"

namespace dgops
{
  namespace
  {
    // VTK's index of lattice point (i, j, l) of a Lagrange quadrilateral
    // (dim 2, l = 0) or hexahedron of order q: corners counter-clockwise
    // per layer, then the points of the edges, of the faces, of the
    // interior (VTK_LAGRANGE_QUADRILATERAL / _HEXAHEDRON node order)
    int
    vtk_lagrange_index(int dim, int q, int i, int j, int l)
    {
      const bool ib = i == 0 || i == q, jb = j == 0 || j == q;
      const bool lb = dim == 3 && (l == 0 || l == q);
      const int  nb = ib + jb + lb;
      const int  m  = q - 1; // interior points per edge
      if (nb == dim)
        return (i ? (j ? 2 : 1) : (j ? 3 : 0)) + (l ? 4 : 0);
      int off = dim == 3 ? 8 : 4;
      if (nb == dim - 1)
        {
          // an edge: along i, along j, or (3D) along l
          if (!ib)
            return off + (i - 1) + (j ? 2 * m : 0) + (l ? 4 * m : 0);
          if (!jb)
            return off + (j - 1) + (i ? m : 3 * m) + (l ? 4 * m : 0);
          off += 8 * m;
          return off + (l - 1) + m * (i ? (j ? 2 : 1) : (j ? 3 : 0));
        }
      off += (dim == 3 ? 12 : 4) * m;
      if (dim == 3 && nb == 1)
        {
          if (ib)
            return off + (j - 1) + m * (l - 1) + (i ? m * m : 0);
          off += 2 * m * m;
          if (jb)
            return off + (i - 1) + m * (l - 1) + (j ? m * m : 0);
          off += 2 * m * m;
          return off + (i - 1) + m * (j - 1) + (l ? m * m : 0);
        }
      if (dim == 3)
        off += 6 * m * m;
      return off + (i - 1) + m * ((j - 1) + m * (dim == 3 ? l - 1 : 0));
    }

    // position of reference point xi in a cell with corners X[v][dim]
    // (corner v at (v&1, v>>1&1, v>>2&1))
    void
    cell_point(int dim, const double *X, const double *xi, double *x)
    {
      const int nv = 1 << dim;
      for (int e = 0; e < 3; ++e)
        x[e] = 0;
      for (int v = 0; v < nv; ++v)
        {
          double w = 1;
          for (int d = 0; d < dim; ++d)
            w *= (v >> d & 1) ? xi[d] : 1 - xi[d];
          for (int e = 0; e < dim; ++e)
            x[e] += w * X[v * dim + e];
        }
    }

    // out[a-lattice] = sum over nodes E[a][i] ... in[i...] (tensor product
    // of the 1D matrix E over dim directions, lexicographic, x fastest)
    void
    tensor_apply(int dim, int n, int m, const std::vector<double> &E,
                 const double *in, double *out)
    {
      std::vector<double> t1(m * n * n), t2(m * m * n);
      if (dim == 2)
        {
          for (int j = 0; j < n; ++j)
            for (int a = 0; a < m; ++a)
              {
                double s = 0;
                for (int i = 0; i < n; ++i)
                  s += E[a * n + i] * in[i + n * j];
                t1[a + m * j] = s;
              }
          for (int b = 0; b < m; ++b)
            for (int a = 0; a < m; ++a)
              {
                double s = 0;
                for (int j = 0; j < n; ++j)
                  s += E[b * n + j] * t1[a + m * j];
                out[a + m * b] = s;
              }
          return;
        }
      for (int l = 0; l < n; ++l)
        for (int j = 0; j < n; ++j)
          for (int a = 0; a < m; ++a)
            {
              double s = 0;
              for (int i = 0; i < n; ++i)
                s += E[a * n + i] * in[i + n * (j + n * l)];
              t1[a + m * (j + n * l)] = s;
            }
      for (int l = 0; l < n; ++l)
        for (int b = 0; b < m; ++b)
          for (int a = 0; a < m; ++a)
            {
              double s = 0;
              for (int j = 0; j < n; ++j)
                s += E[b * n + j] * t1[a + m * (j + n * l)];
              t2[a + m * (b + m * l)] = s;
            }
      for (int c = 0; c < m; ++c)
        for (int b = 0; b < m; ++b)
          for (int a = 0; a < m; ++a)
            {
              double s = 0;
              for (int l = 0; l < n; ++l)
                s += E[c * n + l] * t2[a + m * (b + m * l)];
              out[a + m * (b + m * c)] = s;
            }
    }

    struct Appended
    {
      std::string xml, data;
      void
      array(const char *type, const char *name, int comps, const void *p,
            uint64_t bytes)
      {
        char buf[256];
        std::snprintf(buf, sizeof(buf),
                      "<DataArray type=\"%s\"%s%s%s NumberOfComponents=\"%d\" "
                      "format=\"appended\" offset=\"%llu\"/>\n",
                      type, name ? " Name=\"" : "", name ? name : "",
                      name ? "\"" : "", comps,
                      (unsigned long long)data.size());
        xml += buf;
        data.append(reinterpret_cast<const char *>(&bytes), sizeof(bytes));
        data.append(static_cast<const char *>(p), bytes);
      }
    };
  } // namespace

  void
  write_vtu_host(const dg_config &cfg, const StructuredMesh &mesh,
                 const double *values, double time, const char *filename)
  {
    const int dim = cfg.dim, k = cfg.degree, n = k + 1, nc = cfg.n_components;
    if (k < 1)
      throw Error(DG_ERR_UNSUPPORTED, "VTU output needs degree >= 1");
    const int     nd   = dim == 3 ? n * n * n : n * n; // nodes and points
    const int64_t nce  = mesh.n_cells;
    const int     nv   = 1 << dim;
    const Basis1D basis = make_basis(n);

    // E[a][i] = ell_i(a / k): Gauss-Lobatto Lagrange basis at the
    // equidistant patch points
    std::vector<double> E(n * n);
    for (int a = 0; a < n; ++a)
      for (int i = 0; i < n; ++i)
        {
          const double x = (double)a / k;
          double       v = 1;
          for (int m = 0; m < n; ++m)
            if (m != i)
              v *= (x - basis.gl_pts[m]) / (basis.gl_pts[i] - basis.gl_pts[m]);
          E[a * n + i] = v;
        }
    // lattice point of each VTK node
    std::vector<int> vtk_to_lex(nd);
    for (int l = 0; l < (dim == 3 ? n : 1); ++l)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
          vtk_to_lex[vtk_lagrange_index(dim, k, i, j, l)] =
            i + n * (j + n * l);

    const int64_t       np = nce * nd;
    std::vector<double> pts(3 * np), sol((size_t)nc * np), ana((size_t)nc * np);
    std::vector<int64_t> conn(np), offs(nce);
    std::vector<uint8_t> types(nce, dim == 3 ? 72 : 70);
    std::vector<double>  nodal(nd), exact(nd), at(nd);
    const double        *a = cfg.velocity;
    for (int64_t c = 0; c < nce; ++c)
      {
        const double *X = &mesh.cell_vertices[c * nv * dim];
        for (int p = 0; p < nd; ++p)
          {
            const double xi[3] = {(double)(p % n) / k, (double)(p / n % n) / k,
                                  dim == 3 ? (double)(p / (n * n)) / k : 0.};
            cell_point(dim, X, xi, &pts[3 * (c * nd + p)]);
            conn[c * nd + p] = c * nd + vtk_to_lex[p];
          }
        offs[c] = (c + 1) * nd;
        for (int comp = 0; comp < nc; ++comp)
          {
            const double *u = values + (c * nc + comp) * nd;
            tensor_apply(dim, n, n, E, u, at.data());
            for (int p = 0; p < nd; ++p)
              sol[(size_t)comp * np + c * nd + p] = at[p];
            // the exact solution interpolated at the nodes, then evaluated
            // like the solution
            for (int p = 0; p < nd; ++p)
              {
                const double xi[3] = {basis.gl_pts[p % n],
                                      basis.gl_pts[p / n % n],
                                      dim == 3 ? basis.gl_pts[p / (n * n)] : 0.};
                double       x[3];
                cell_point(dim, X, xi, x);
                exact[p] = dim == 3
                             ? analytic_solution<3>(cfg.solution_kind,
                                                    cfg.solution_param, time,
                                                    a, x, comp)
                             : analytic_solution<2>(cfg.solution_kind,
                                                    cfg.solution_param, time,
                                                    a, x, comp);
              }
            tensor_apply(dim, n, n, E, exact.data(), at.data());
            for (int p = 0; p < nd; ++p)
              ana[(size_t)comp * np + c * nd + p] = at[p];
          }
      }

    Appended    A;
    std::string head;
    char        buf[512];
    std::snprintf(buf, sizeof(buf),
                  "<?xml version=\"1.0\"?>\n"
                  "<VTKFile type=\"UnstructuredGrid\" version=\"2.2\" "
                  "byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
                  "<UnstructuredGrid>\n"
                  "<FieldData>\n");
    head += buf;
    A.xml.clear();
    A.array("Float64", "TIME", 1, &time, sizeof(double));
    head += A.xml;
    std::snprintf(buf, sizeof(buf),
                  "</FieldData>\n<Piece NumberOfPoints=\"%lld\" "
                  "NumberOfCells=\"%lld\">\n<PointData>\n",
                  (long long)np, (long long)nce);
    head += buf;
    for (int comp = 0; comp < nc; ++comp)
      {
        const std::string sfx = nc > 1 ? "_" + std::to_string(comp) : "";
        A.xml.clear();
        A.array("Float64", ("solution" + sfx).c_str(), 1,
                sol.data() + (size_t)comp * np, sizeof(double) * np);
        A.array("Float64", ("analytic_solution" + sfx).c_str(), 1,
                ana.data() + (size_t)comp * np, sizeof(double) * np);
        head += A.xml;
      }
    head += "</PointData>\n<Points>\n";
    A.xml.clear();
    A.array("Float64", nullptr, 3, pts.data(), sizeof(double) * 3 * np);
    head += A.xml;
    head += "</Points>\n<Cells>\n";
    A.xml.clear();
    A.array("Int64", "connectivity", 1, conn.data(), sizeof(int64_t) * np);
    A.array("Int64", "offsets", 1, offs.data(), sizeof(int64_t) * nce);
    A.array("UInt8", "types", 1, types.data(), nce);
    head += A.xml;
    head += "</Cells>\n</Piece>\n</UnstructuredGrid>\n"
            "<AppendedData encoding=\"raw\">\n_";
    FILE *f = std::fopen(filename, "wb");
    if (!f)
      throw Error(DG_ERR_ARGUMENT, std::string("cannot open ") + filename);
    bool ok = std::fwrite(head.data(), 1, head.size(), f) == head.size() &&
              std::fwrite(A.data.data(), 1, A.data.size(), f) ==
                A.data.size();
    const char *tail = "\n</AppendedData>\n</VTKFile>\n";
    ok = ok && std::fwrite(tail, 1, std::strlen(tail), f) == std::strlen(tail);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok)
      throw Error(DG_ERR_ARGUMENT, std::string("write failed: ") + filename);
  }
} // namespace dgops
