// CPU oracle (test infrastructure, not part of the product): a plain,
// face-centric restatement of the reference's matrix-free DG operators,
// following deal.II's FEEvaluation / FEFaceEvaluation semantics:
//   gather_evaluate   -> values and unit-cell gradients at Gauss points by
//                        sum factorization with the shape value matrix
//                        S[q][i] = phi_i(x_q) and the shape gradient matrix
//                        G[q][i] = phi_i'(x_q)
//   get_gradient      -> J^-T grad_ref
//   submit_gradient F -> J^-1 F JxW,  submit_value v -> v JxW
//   integrate_scatter -> transposed sweeps, added into dst
//   faces             -> one loop over faces, minus/plus evaluators with the
//                        minus side's unit normal, both sides integrated
//   inverse mass      -> separate cell loop, S^-1 diag(1/JxW) S^-T
// Geometry is evaluated per point from the cell's corners with the
// trilinear shape functions (MappingQ of a straight-sided mesh).
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <vector>

#ifdef _OPENMP
#  include <omp.h>
#endif

namespace
{
  constexpr double PI = 3.14159265358979323846;
  using ld            = long double;

  // ---- 1D polynomial data ----------------------------------------------------

  void
  legendre(int m, ld x, ld &p, ld &dp)
  {
    ld p0 = 1, p1 = x;
    if (m == 0)
      {
        p  = 1;
        dp = 0;
        return;
      }
    for (int k = 2; k <= m; ++k)
      {
        ld pk = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0    = p1;
        p1    = pk;
      }
    p  = p1;
    dp = m * (x * p1 - p0) / (x * x - 1);
  }

  // Gauss points on [0,1] by Newton on P_n
  void
  gauss(int n, std::vector<ld> &x, std::vector<ld> &w)
  {
    x.resize(n);
    w.resize(n);
    for (int i = 0; i < n; ++i)
      {
        ld t = std::cos(PI * (n - i - 0.25L) / (n + 0.5L));
        for (int it = 0; it < 200; ++it)
          {
            ld p, dp;
            legendre(n, t, p, dp);
            ld dt = p / dp;
            t -= dt;
            if (std::fabs((double)dt) < 1e-20)
              break;
          }
        ld p, dp;
        legendre(n, t, p, dp);
        x[i] = (1 + t) / 2;
        w[i] = 1 / ((1 - t * t) * dp * dp);
      }
  }

  // Gauss-Lobatto nodes on [0,1]: 0, 1 and the extrema of P_{n-1}, found by
  // bisection on P'_{n-1} between consecutive Gauss points of P_{n-1}
  std::vector<ld>
  lobatto(int n)
  {
    std::vector<ld> x(n);
    x[0]     = 0;
    x[n - 1] = 1;
    if (n <= 2)
      return x;
    std::vector<ld> g, w;
    gauss(n - 1, g, w); // roots of P_{n-1} on [0,1], interlace the extrema
    for (int i = 1; i < n - 1; ++i)
      {
        ld a = 2 * g[i - 1] - 1, b = 2 * g[i] - 1;
        ld pa, dpa, pb, dpb;
        legendre(n - 1, a, pa, dpa);
        for (int it = 0; it < 200; ++it)
          {
            ld m = (a + b) / 2, pm, dpm;
            legendre(n - 1, m, pm, dpm);
            if ((dpm > 0) == (dpa > 0))
              {
                a   = m;
                dpa = dpm;
              }
            else
              b = m;
          }
        x[i] = (1 + (a + b) / 2) / 2;
      }
    return x;
  }

  ld
  lagr(const std::vector<ld> &nodes, int i, ld x)
  {
    ld v = 1;
    for (size_t j = 0; j < nodes.size(); ++j)
      if ((int)j != i)
        v *= (x - nodes[j]) / (nodes[i] - nodes[j]);
    return v;
  }

  ld
  lagr_der(const std::vector<ld> &nodes, int i, ld x)
  {
    // derivative by the product rule, summing over the omitted factor
    ld s = 0;
    for (size_t k = 0; k < nodes.size(); ++k)
      {
        if ((int)k == i)
          continue;
        ld p = 1 / (nodes[i] - nodes[k]);
        for (size_t j = 0; j < nodes.size(); ++j)
          if ((int)j != i && j != k)
            p *= (x - nodes[j]) / (nodes[i] - nodes[j]);
        s += p;
      }
    return s;
  }

  // inverse of a dense n x n matrix by Gauss-Jordan with partial pivoting
  std::vector<double>
  invert(std::vector<double> a, int n)
  {
    std::vector<double> inv(n * n, 0.);
    for (int i = 0; i < n; ++i)
      inv[i * n + i] = 1;
    for (int c = 0; c < n; ++c)
      {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
          if (std::fabs(a[r * n + c]) > std::fabs(a[piv * n + c]))
            piv = r;
        for (int k = 0; k < n; ++k)
          {
            std::swap(a[c * n + k], a[piv * n + k]);
            std::swap(inv[c * n + k], inv[piv * n + k]);
          }
        const double d = a[c * n + c];
        for (int k = 0; k < n; ++k)
          {
            a[c * n + k] /= d;
            inv[c * n + k] /= d;
          }
        for (int r = 0; r < n; ++r)
          if (r != c)
            {
              const double f = a[r * n + c];
              if (f == 0)
                continue;
              for (int k = 0; k < n; ++k)
                {
                  a[r * n + k] -= f * a[c * n + k];
                  inv[r * n + k] -= f * inv[c * n + k];
                }
            }
      }
    return inv;
  }

  // ---- analytic data -------------------------------------------------------------

  void
  velocity(const or_problem &P, const double *x, double tf, double *a)
  {
    const int dim = P.dim;
    for (int d = 0; d < 3; ++d)
      a[d] = 0;
    switch (P.velocity_kind)
      {
        case 1: // sin2πy cos2πz, sin2πz cos2πx, sin2πx cos2πy
          a[0] = std::sin(2 * PI * x[1]) * std::cos(2 * PI * x[2]);
          a[1] = std::sin(2 * PI * x[2]) * std::cos(2 * PI * x[0]);
          a[2] = std::sin(2 * PI * x[0]) * std::cos(2 * PI * x[1]);
          break;
        case 2: // advection_solver_variable.cc:163-176
        case 3: // advection_solver_implicit_real.cc:168-184
          {
            const double factor = 2.;
            a[0] = factor * (std::sin(2 * PI * x[1]) * std::sin(PI * x[0]) *
                             std::sin(PI * x[0]));
            a[1] = -factor * (std::sin(2 * PI * x[0]) * std::sin(PI * x[1]) *
                              std::sin(PI * x[1]));
            if (P.velocity_kind == 3)
              {
                a[0] += factor * 0.2 * std::sin(80 * PI * (x[0] + 0.2)) *
                        std::cos(80 * PI * (x[1] + 0.3));
                a[1] -= factor * 0.2 * std::cos(80 * PI * (x[0] + 0.2)) *
                        std::sin(80 * PI * (x[1] + 0.3));
              }
            break;
          }
        default:
          for (int d = 0; d < dim; ++d)
            a[d] = P.velocity[d];
      }
    for (int d = 0; d < dim; ++d)
      a[d] *= tf;
  }

  double
  solution(const or_problem &P, const double *x, double t, int comp = 0)
  {
    const int     dim = P.dim;
    const double *p   = P.solution_param;
    const double *a   = P.velocity;
    switch (P.solution_kind)
      {
        case 1:
          {
            double r = std::sin(4. * PI * (x[0] - t * a[0]));
            for (int d = 1; d < dim; ++d)
              r *= std::cos(4. * PI * (x[d] - t * a[d]));
            return r;
          }
        case 2:
          {
            double r = 1;
            for (int d = 0; d < dim; ++d)
              r *= std::sin(p[0] * (x[d] - t * a[d]));
            if (p[1] != 0)
              r *= std::exp(-p[1] * dim * p[0] * p[0] * t);
            return r;
          }
        case 3:
          {
            double r2 = 0;
            for (int d = 0; d < dim; ++d)
              r2 += (x[d] - p[d]) * (x[d] - p[d]);
            return std::exp(-r2 / (p[3] * p[3]));
          }
        case 4:
          {
            double s = 1;
            for (int d = 0; d < dim; ++d)
              s += (d + 1) * (x[d] - t * a[d]);
            return std::pow(s, (int)p[0]);
          }
        case 5:
          return std::exp(-400. * ((x[0] - 0.5) * (x[0] - 0.5) +
                                   (x[1] - 0.75) * (x[1] - 0.75))) +
                 std::exp(-100. * ((x[0] - 0.5) * (x[0] - 0.5) +
                                   (x[1] - 0.25) * (x[1] - 0.25)));
        case 6:
          return comp == 0 ? -std::sin(PI * x[0]) : 0.;
        default:
          return 0;
      }
  }
} // namespace

struct or_handle
{
  or_problem P;
  int        dim, n, nd, nq1;
  int        N[3];
  double     h[3];
  int64_t    n_cells;
  std::vector<int>    ijk, nbr;
  std::vector<double> corners; // [cell][2^dim][dim]
  double              min_dist;
  // 1D data: the operator's quadrature (QGauss(n), or QGaussLobatto(n)
  // under use_gl_quad) and the mass matrix's (xqm, wqm, Sm, Sminv: the
  // operator's, or QGaussLobatto(n) under use_gl_quad_mass)
  std::vector<double> xq, wq, S, G, Sinv, phi0, phi1;
  std::vector<double> xqm, wqm, Sm, Sminv;
  std::vector<ld>     xgl;

  double time_factor(double t) const
  {
    return P.time_factor == 1 ? std::cos(PI * t / P.final_time) : 1.;
  }
};

namespace
{
  // ---- mesh -------------------------------------------------------------------

  int64_t
  cell_of(const or_handle &H, int i, int j, int k)
  {
    // coarse cells lexicographic; inside, the child tree of refine_global
    const int r      = H.P.refine;
    int64_t   coarse = (i >> r) + (int64_t)H.P.n_coarse[0] *
                                   ((j >> r) + (int64_t)H.P.n_coarse[1] * (k >> r));
    int64_t   idx    = coarse;
    for (int level = r - 1; level >= 0; --level)
      {
        int child = ((i >> level) & 1) | (((j >> level) & 1) << 1);
        if (H.dim == 3)
          child |= ((k >> level) & 1) << 2;
        idx = idx * (1 << H.dim) + child;
      }
    return idx;
  }

  void
  build_mesh(or_handle &H)
  {
    const or_problem &P = H.P;
    H.n_cells           = 1;
    for (int d = 0; d < 3; ++d)
      {
        H.N[d] = d < H.dim ? (P.n_coarse[d] << P.refine) : 1;
        H.h[d] = d < H.dim ? (P.hi[d] - P.lo[d]) / H.N[d] : 1;
        H.n_cells *= H.N[d];
      }
    H.ijk.assign(3 * H.n_cells, 0);
    for (int k = 0; k < H.N[2]; ++k)
      for (int j = 0; j < H.N[1]; ++j)
        for (int i = 0; i < H.N[0]; ++i)
          {
            const int64_t c = cell_of(H, i, j, k);
            H.ijk[3 * c]     = i;
            H.ijk[3 * c + 1] = j;
            H.ijk[3 * c + 2] = k;
          }
    const int V[3] = {H.N[0] + 1, H.N[1] + 1, H.dim == 3 ? H.N[2] + 1 : 1};
    std::vector<double> vert((int64_t)V[0] * V[1] * V[2] * H.dim);
    std::mt19937_64     rng(P.seed);
    for (int k = 0; k < V[2]; ++k)
      for (int j = 0; j < V[1]; ++j)
        for (int i = 0; i < V[0]; ++i)
          {
            const int     id[3] = {i, j, k};
            const int64_t v     = i + (int64_t)V[0] * (j + (int64_t)V[1] * k);
            bool          inner = true;
            for (int d = 0; d < H.dim; ++d)
              {
                if (id[d] == 0 || id[d] == H.N[d])
                  inner = false;
                vert[v * H.dim + d] =
                  id[d] == H.N[d] ? P.hi[d] :
                                    P.lo[d] + (P.hi[d] - P.lo[d]) * id[d] / H.N[d];
              }
            if (inner && P.perturbation != 0)
              for (int d = 0; d < H.dim; ++d)
                {
                  const double u = (double)(rng() >> 11) * 0x1.0p-53;
                  vert[v * H.dim + d] += P.perturbation * H.h[d] * (2 * u - 1);
                }
          }
    const int nv = 1 << H.dim;
    H.corners.resize(H.n_cells * nv * H.dim);
    H.nbr.resize(H.n_cells * 2 * H.dim);
    H.min_dist = 1e300;
    for (int64_t c = 0; c < H.n_cells; ++c)
      {
        const int *id = &H.ijk[3 * c];
        for (int v = 0; v < nv; ++v)
          {
            const int64_t g = (id[0] + (v & 1)) +
                              (int64_t)V[0] * ((id[1] + ((v >> 1) & 1)) +
                                               (int64_t)V[1] * (id[2] + ((v >> 2) & 1)));
            for (int d = 0; d < H.dim; ++d)
              H.corners[(c * nv + v) * H.dim + d] = vert[g * H.dim + d];
          }
        for (int a = 0; a < nv; ++a)
          for (int b = a + 1; b < nv; ++b)
            {
              double s = 0;
              for (int d = 0; d < H.dim; ++d)
                {
                  const double t = H.corners[(c * nv + a) * H.dim + d] -
                                   H.corners[(c * nv + b) * H.dim + d];
                  s += t * t;
                }
              H.min_dist = std::min(H.min_dist, std::sqrt(s));
            }
        for (int d = 0; d < H.dim; ++d)
          for (int s = 0; s < 2; ++s)
            {
              int q[3] = {id[0], id[1], id[2]};
              q[d] += s ? 1 : -1;
              int64_t nb = -1;
              if (q[d] < 0 || q[d] >= H.N[d])
                {
                  if (P.periodic[d])
                    {
                      q[d] = (q[d] + H.N[d]) % H.N[d];
                      nb   = cell_of(H, q[0], q[1], q[2]);
                    }
                }
              else
                nb = cell_of(H, q[0], q[1], q[2]);
              H.nbr[c * 2 * H.dim + 2 * d + s] = (int)nb;
            }
      }
  }

  // ---- geometry ------------------------------------------------------------------

  // trilinear shape function N_v and its gradient at xi
  void
  map_point(const or_handle &H, int64_t cell, const double *xi, double *x,
            double J[3][3])
  {
    const int     nv = 1 << H.dim;
    const double *X  = &H.corners[cell * nv * H.dim];
    for (int e = 0; e < 3; ++e)
      {
        x[e] = 0;
        for (int d = 0; d < 3; ++d)
          J[e][d] = 0;
      }
    for (int v = 0; v < nv; ++v)
      {
        double N = 1, dN[3];
        for (int d = 0; d < H.dim; ++d)
          {
            const int b = (v >> d) & 1;
            N *= b ? xi[d] : 1 - xi[d];
          }
        for (int d = 0; d < H.dim; ++d)
          {
            dN[d] = 1;
            for (int e = 0; e < H.dim; ++e)
              {
                const int b = (v >> e) & 1;
                if (e == d)
                  dN[d] *= b ? 1. : -1.;
                else
                  dN[d] *= b ? xi[e] : 1 - xi[e];
              }
          }
        for (int e = 0; e < H.dim; ++e)
          {
            x[e] += N * X[v * H.dim + e];
            for (int d = 0; d < H.dim; ++d)
              J[e][d] += dN[d] * X[v * H.dim + e];
          }
      }
  }

  double
  inverse_jacobian(int dim, const double J[3][3], double Ji[3][3])
  {
    if (dim == 2)
      {
        const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        Ji[0][0]         = J[1][1] / det;
        Ji[0][1]         = -J[0][1] / det;
        Ji[1][0]         = -J[1][0] / det;
        Ji[1][1]         = J[0][0] / det;
        return det;
      }
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
    Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
    Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
    Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
    Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    return det;
  }

  // ---- sum factorization on a cell ------------------------------------------------

  // out = (M_{dim-1} x ... x M_0) in with M_d an (m x k) matrix per direction
  // stored row-major; extents k^dim -> m^dim. trans: use M^T (k x m).
  void
  tensor_apply(int dim, const double *const *M, int m, int k, bool trans,
               const double *in, double *out, std::vector<double> &tmp1,
               std::vector<double> &tmp2)
  {
    int ext[3] = {k, k, dim == 3 ? k : 1};
    tmp1.assign(in, in + ext[0] * ext[1] * ext[2]);
    for (int d = 0; d < dim; ++d)
      {
        int eo[3] = {ext[0], ext[1], ext[2]};
        eo[d]     = m;
        tmp2.assign(eo[0] * eo[1] * eo[2], 0.);
        for (int c = 0; c < eo[2]; ++c)
          for (int b = 0; b < eo[1]; ++b)
            for (int a = 0; a < eo[0]; ++a)
              {
                int    q[3] = {a, b, c};
                const int qo = q[d];
                double s     = 0;
                for (int i = 0; i < ext[d]; ++i)
                  {
                    q[d]           = i;
                    const double mv = trans ? M[d][i * m + qo] : M[d][qo * k + i];
                    s += mv * tmp1[q[0] + ext[0] * (q[1] + ext[1] * q[2])];
                  }
                tmp2[a + eo[0] * (b + eo[1] * c)] = s;
              }
        ext[d] = m;
        tmp1.swap(tmp2);
      }
    std::copy(tmp1.begin(), tmp1.end(), out);
  }

  struct CellEval
  {
    std::vector<double> vals, grads[3], tmp1, tmp2, buf;
  };

  // values and unit-cell gradients at the Gauss points of a cell
  void
  evaluate(const or_handle &H, const double *u, CellEval &E, bool grads)
  {
    const int     n  = H.n;
    const double *Sm = H.S.data(), *Gm = H.G.data();
    const double *M[3] = {Sm, Sm, Sm};
    E.vals.resize(H.nd);
    tensor_apply(H.dim, M, n, n, false, u, E.vals.data(), E.tmp1, E.tmp2);
    if (grads)
      for (int d = 0; d < H.dim; ++d)
        {
          const double *Md[3] = {Sm, Sm, Sm};
          Md[d]               = Gm;
          E.grads[d].resize(H.nd);
          tensor_apply(H.dim, Md, n, n, false, u, E.grads[d].data(), E.tmp1,
                       E.tmp2);
        }
  }

  // dst += S^T vals + sum_d (G in d)^T grads[d]
  void
  integrate_add(const or_handle &H, const double *vals,
                const std::vector<double> *grads, double *dst, CellEval &E)
  {
    const int     n  = H.n;
    const double *Sm = H.S.data(), *Gm = H.G.data();
    E.buf.resize(H.nd);
    if (vals)
      {
        const double *M[3] = {Sm, Sm, Sm};
        tensor_apply(H.dim, M, n, n, true, vals, E.buf.data(), E.tmp1,
                     E.tmp2);
        for (int i = 0; i < H.nd; ++i)
          dst[i] += E.buf[i];
      }
    if (grads)
      for (int d = 0; d < H.dim; ++d)
        {
          const double *M[3] = {Sm, Sm, Sm};
          M[d]               = Gm;
          tensor_apply(H.dim, M, n, n, true, grads[d].data(), E.buf.data(),
                       E.tmp1, E.tmp2);
          for (int i = 0; i < H.nd; ++i)
            dst[i] += E.buf[i];
        }
  }

  void
  q_index(const or_handle &H, int q, int *qi)
  {
    qi[0] = q % H.n;
    qi[1] = (q / H.n) % H.n;
    qi[2] = H.dim == 3 ? q / (H.n * H.n) : 0;
  }

  // ---- advection cell and face terms ---------------------------------------------

  // Which operator the loops evaluate. Explicit files submit the terms with
  // the signs of advection_solver.cc:498-508 and :560-575; the implicit
  // files the opposite signs plus inv_dt u (advection_solver_implicit_real.cc
  // :1003-1011, :1043-1055).
  struct LoopSpec
  {
    bool   implicit;
    double inv_dt;
    int    n_stages;              // stage-coupled IRK system when > 1
    double time_factor[8];        // per stage
    std::vector<double> mass;     // stage coupling [c][b] (IRK), size s*s
  };

  // cell term for one cell and (possibly) several stage blocks
  void
  cell_term(const or_handle &H, const LoopSpec &L, int64_t cell,
            const double *src, int64_t block, double *dst, CellEval &E)
  {
    const int dim = H.dim, n = H.n, nd = H.nd;
    const double s  = H.P.factor_skew;
    const double nu = H.P.operator_kind == 1 ? H.P.viscosity : 0.;
    const int    ns = L.n_stages;
    std::vector<CellEval> ev(ns);
    for (int c = 0; c < ns; ++c)
      evaluate(H, src + c * block + cell * nd, ev[c], true);
    for (int c = 0; c < ns; ++c)
      {
        std::vector<double> vals(nd), grads[3];
        for (int d = 0; d < dim; ++d)
          grads[d].assign(nd, 0.);
        for (int q = 0; q < nd; ++q)
          {
            int qi[3];
            q_index(H, q, qi);
            double xi[3] = {H.xq[qi[0]], H.xq[qi[1]], dim == 3 ? H.xq[qi[2]] : 0};
            double x[3], J[3][3], Ji[3][3], a[3];
            map_point(H, cell, xi, x, J);
            const double det = inverse_jacobian(dim, J, Ji);
            double       w   = H.wq[qi[0]] * H.wq[qi[1]];
            if (dim == 3)
              w *= H.wq[qi[2]];
            const double JxW = det * w;
            velocity(H.P, x, L.time_factor[c], a);
            const double u = ev[c].vals[q];
            // real-space gradient J^-T grad_ref
            double gu[3] = {0, 0, 0};
            for (int e = 0; e < dim; ++e)
              for (int d = 0; d < dim; ++d)
                gu[e] += Ji[d][e] * ev[c].grads[d][q];
            double adotgu = 0;
            for (int e = 0; e < dim; ++e)
              adotgu += a[e] * gu[e];
            double F[3], V;
            if (!L.implicit)
              {
                for (int e = 0; e < dim; ++e)
                  F[e] = (1.0 - s) * a[e] * u - nu * gu[e];
                V = -s * adotgu;
              }
            else
              {
                // advection_diffusion_solver.cc:546-549: volume_flux + nu grad u
                for (int e = 0; e < dim; ++e)
                  F[e] = (-1.0 + s) * a[e] * u + nu * gu[e];
                V = s * adotgu;
                if (ns == 1)
                  V += L.inv_dt * u;
                else
                  for (int b = 0; b < ns; ++b)
                    V += L.mass[c * ns + b] * ev[b].vals[q];
              }
            // submit_gradient: J^-1 F JxW; submit_value: V JxW
            for (int d = 0; d < dim; ++d)
              {
                double t = 0;
                for (int e = 0; e < dim; ++e)
                  t += Ji[d][e] * F[e];
                grads[d][q] = t * JxW;
              }
            vals[q] = V * JxW;
          }
        integrate_add(H, vals.data(), grads, dst + c * block + cell * nd, E);
      }
  }

  // face values: trace of a cell on face (d, side) at the face Gauss points;
  // integrate: transposed
  void
  face_eval(const or_handle &H, const double *u, int d, int side, double *fv)
  {
    const int     n = H.n, dim = H.dim;
    const double *tr = side ? H.phi1.data() : H.phi0.data();
    const int     nf = dim == 3 ? n * n : n;
    for (int fq = 0; fq < nf; ++fq)
      {
        int a = fq % n, b = dim == 3 ? fq / n : 0;
        double s = 0;
        for (int k = 0; k < (dim == 3 ? n : 1); ++k)
          for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i)
              {
                int node[3] = {i, j, k};
                // tangential directions ascending
                int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
                double f = tr[node[d]] * H.S[a * n + node[t1]];
                if (dim == 3)
                  f *= H.S[b * n + node[t2]];
                s += f * u[i + n * (j + n * k)];
              }
        fv[fq] = s;
      }
  }

  void
  face_integrate_add(const or_handle &H, const double *fv, int d, int side,
                     double *dst)
  {
    const int     n = H.n, dim = H.dim;
    const double *tr = side ? H.phi1.data() : H.phi0.data();
    const int     nf = dim == 3 ? n * n : n;
    for (int k = 0; k < (dim == 3 ? n : 1); ++k)
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i)
          {
            int node[3] = {i, j, k};
            int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
            double s = 0;
            for (int fq = 0; fq < nf; ++fq)
              {
                int a = fq % n, b = dim == 3 ? fq / n : 0;
                double f = tr[node[d]] * H.S[a * n + node[t1]];
                if (dim == 3)
                  f *= H.S[b * n + node[t2]];
                s += f * fv[fq];
              }
            dst[i + n * (j + n * k)] += s;
          }
  }

  // outward unit normal and JxW at a face point of `cell` on face (d,side)
  void
  face_geometry(const or_handle &H, int64_t cell, int d, int side,
                const int *tq, double *x, double *normal, double &JxW)
  {
    const int dim = H.dim;
    double    xi[3] = {0, 0, 0};
    int       t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    xi[d]  = side;
    xi[t1] = H.xq[tq[0]];
    double w = H.wq[tq[0]];
    if (dim == 3)
      {
        xi[t2] = H.xq[tq[1]];
        w *= H.wq[tq[1]];
      }
    double J[3][3];
    map_point(H, cell, xi, x, J);
    double nv[3] = {0, 0, 0};
    if (dim == 3)
      {
        // area vector of the face, oriented along +xi_d
        const int a = (d + 1) % 3, b = (d + 2) % 3;
        nv[0] = J[1][a] * J[2][b] - J[2][a] * J[1][b];
        nv[1] = J[2][a] * J[0][b] - J[0][a] * J[2][b];
        nv[2] = J[0][a] * J[1][b] - J[1][a] * J[0][b];
      }
    else
      {
        const int t = 1 - d;
        // tangent J[.][t]; rotate to the +xi_d side
        nv[0] = d == 0 ? J[1][t] : -J[1][t];
        nv[1] = d == 0 ? -J[0][t] : J[0][t];
      }
    double len = 0;
    for (int e = 0; e < dim; ++e)
      len += nv[e] * nv[e];
    len = std::sqrt(len);
    const double sg = side ? 1. : -1.;
    for (int e = 0; e < dim; ++e)
      normal[e] = sg * nv[e] / len;
    JxW = len * w;
  }

  // all faces; `block` is the stage block size for IRK
  void
  face_terms(const or_handle &H, const LoopSpec &L, const double *src,
             int64_t block, double *dst, double t_bc)
  {
    const int dim = H.dim, n = H.n, nd = H.nd;
    const int nf  = dim == 3 ? n * n : n;
    const double alpha = H.P.flux_alpha, s = H.P.factor_skew;
    for (int d = 0; d < dim; ++d)
      for (int pass = 0; pass < 3; ++pass)
        {
#pragma omp parallel for schedule(static)
          for (int64_t cell = 0; cell < H.n_cells; ++cell)
            {
              const int id = H.ijk[3 * cell + d];
              const int mypass = (H.N[d] % 2 == 1 && id == H.N[d] - 1) ? 2 : id % 2;
              if (mypass != pass)
                continue;
              // interior / boundary faces on the upper side, boundary faces
              // on the lower side
              for (int side = 0; side < 2; ++side)
                {
                  const int64_t nb = H.nbr[cell * 2 * dim + 2 * d + side];
                  if (side == 0 && nb >= 0)
                    continue; // counted by the neighbour's upper face
                  std::vector<double> um(nf), up(nf), fm(nf), fp(nf);
                  for (int c = 0; c < L.n_stages; ++c)
                    {
                      const double *sc = src + c * block;
                      face_eval(H, sc + cell * nd, d, side, um.data());
                      if (nb >= 0)
                        face_eval(H, sc + nb * nd, d, 1 - side, up.data());
                      for (int fq = 0; fq < nf; ++fq)
                        {
                          int tq[2] = {fq % n, dim == 3 ? fq / n : 0};
                          double x[3], normal[3], JxW, a[3];
                          face_geometry(H, cell, d, side, tq, x, normal, JxW);
                          velocity(H.P, x, L.time_factor[c], a);
                          double an = 0;
                          for (int e = 0; e < dim; ++e)
                            an += a[e] * normal[e];
                          const double uminus = um[fq];
                          const double uplus =
                            nb >= 0 ? up[fq] : solution(H.P, x, t_bc);
                          const double flux =
                            0.5 * ((uminus + uplus) * an +
                                   alpha * std::fabs(an) * (uminus - uplus));
                          if (!L.implicit)
                            {
                              fm[fq] = (-flux + s * an * uminus) * JxW;
                              fp[fq] = (flux - s * an * uplus) * JxW;
                            }
                          else
                            {
                              fm[fq] = (flux - s * an * uminus) * JxW;
                              fp[fq] = (-flux + s * an * uplus) * JxW;
                            }
                        }
                      face_integrate_add(H, fm.data(), d, side,
                                         dst + c * block + cell * nd);
                      if (nb >= 0)
                        face_integrate_add(H, fp.data(), d, 1 - side,
                                           dst + c * block + nb * nd);
                    }
                }
            }
        }
  }

  // value and unit-cell gradient of a cell's nodal expansion at an arbitrary
  // reference point, by direct summation over the Lagrange products
  void
  point_eval(const or_handle &H, const double *u, const double *xi,
             double &val, double *gref)
  {
    const int n = H.n, dim = H.dim;
    double    v1[3][8], d1[3][8];
    for (int e = 0; e < dim; ++e)
      for (int i = 0; i < n; ++i)
        {
          v1[e][i] = (double)lagr(H.xgl, i, xi[e]);
          d1[e][i] = (double)lagr_der(H.xgl, i, xi[e]);
        }
    val = 0;
    for (int e = 0; e < 3; ++e)
      gref[e] = 0;
    for (int idx = 0; idx < H.nd; ++idx)
      {
        const int ii[3] = {idx % n, (idx / n) % n, dim == 3 ? idx / (n * n) : 0};
        double    prod  = 1;
        for (int e = 0; e < dim; ++e)
          prod *= v1[e][ii[e]];
        val += prod * u[idx];
        for (int g = 0; g < dim; ++g)
          {
            double pg = 1;
            for (int e = 0; e < dim; ++e)
              pg *= (e == g ? d1[e][ii[e]] : v1[e][ii[e]]);
            gref[g] += pg * u[idx];
          }
      }
  }

  // dst[i] += value * phi_i(xi) + dn * (n . J^-T grad_ref phi_i)(xi)
  void
  point_integrate(const or_handle &H, const double *xi, const double Ji[3][3],
                  const double *normal, double value, double dn, double *dst)
  {
    const int n = H.n, dim = H.dim;
    double    v1[3][8], d1[3][8];
    for (int e = 0; e < dim; ++e)
      for (int i = 0; i < n; ++i)
        {
          v1[e][i] = (double)lagr(H.xgl, i, xi[e]);
          d1[e][i] = (double)lagr_der(H.xgl, i, xi[e]);
        }
    for (int idx = 0; idx < H.nd; ++idx)
      {
        const int ii[3] = {idx % n, (idx / n) % n, dim == 3 ? idx / (n * n) : 0};
        double    prod  = 1;
        for (int e = 0; e < dim; ++e)
          prod *= v1[e][ii[e]];
        double gref[3] = {0, 0, 0};
        for (int g = 0; g < dim; ++g)
          {
            double pg = 1;
            for (int e = 0; e < dim; ++e)
              pg *= (e == g ? d1[e][ii[e]] : v1[e][ii[e]]);
            gref[g] = pg;
          }
        double ndg = 0;
        for (int e = 0; e < dim; ++e)
          {
            double ge = 0;
            for (int g = 0; g < dim; ++g)
              ge += Ji[g][e] * gref[g];
            ndg += normal[e] * ge;
          }
        dst[idx] += value * prod + dn * ndg;
      }
  }

  // dst[i] += sum_d grad_ref phi_i(xi)[d] g[d]
  void
  point_integrate_grad(const or_handle &H, const double *xi, const double *g,
                       double *dst)
  {
    const int n = H.n, dim = H.dim;
    double    v1[3][8], d1[3][8];
    for (int e = 0; e < dim; ++e)
      for (int i = 0; i < n; ++i)
        {
          v1[e][i] = (double)lagr(H.xgl, i, xi[e]);
          d1[e][i] = (double)lagr_der(H.xgl, i, xi[e]);
        }
    for (int idx = 0; idx < H.nd; ++idx)
      {
        const int ii[3] = {idx % n, (idx / n) % n, dim == 3 ? idx / (n * n) : 0};
        double    s     = 0;
        for (int gd = 0; gd < dim; ++gd)
          {
            double pg = 1;
            for (int e = 0; e < dim; ++e)
              pg *= (e == gd ? d1[e][ii[e]] : v1[e][ii[e]]);
            s += pg * g[gd];
          }
        dst[idx] += s;
      }
  }

  // Burgers loop (burgers_solver.cc:259-341): cell flux u (x) u and Rusanov
  // faces, all with QGauss(k(k+1)/2), explicit signs; dst = -L(u)
  void
  burgers_loop(const or_handle &H, const double *src, double *dst)
  {
    const int dim = H.dim, nd = H.nd, nc = dim;
    const int m   = H.P.degree * (H.P.degree + 1) / 2;
    std::vector<ld> gx, gw;
    gauss(m, gx, gw);
    std::fill(dst, dst + H.n_cells * nc * nd, 0.);
    const int mp = dim == 3 ? m * m * m : m * m;
#pragma omp parallel for schedule(static)
    for (int64_t cell = 0; cell < H.n_cells; ++cell)
      for (int q = 0; q < mp; ++q)
        {
          const int qi[3] = {q % m, (q / m) % m, dim == 3 ? q / (m * m) : 0};
          double    xi[3] = {(double)gx[qi[0]], (double)gx[qi[1]],
                             dim == 3 ? (double)gx[qi[2]] : 0.};
          double    w = (double)(gw[qi[0]] * gw[qi[1]]);
          if (dim == 3)
            w *= (double)gw[qi[2]];
          double x[3], J[3][3], Ji[3][3];
          map_point(H, cell, xi, x, J);
          const double det = inverse_jacobian(dim, J, Ji);
          double       u[3], gr[3];
          for (int c = 0; c < nc; ++c)
            point_eval(H, src + (cell * nc + c) * nd, xi, u[c], gr);
          for (int c = 0; c < nc; ++c)
            {
              double g[3];
              for (int d = 0; d < dim; ++d)
                {
                  double s = 0;
                  for (int e = 0; e < dim; ++e)
                    s += Ji[d][e] * u[c] * u[e];
                  g[d] = s * det * w;
                }
              point_integrate_grad(H, xi, g, dst + (cell * nc + c) * nd);
            }
        }
    const int mf = dim == 3 ? m * m : m;
    for (int d = 0; d < dim; ++d)
      for (int pass = 0; pass < 3; ++pass)
        {
#pragma omp parallel for schedule(static)
          for (int64_t cell = 0; cell < H.n_cells; ++cell)
            {
              const int id = H.ijk[3 * cell + d];
              const int mypass =
                (H.N[d] % 2 == 1 && id == H.N[d] - 1) ? 2 : id % 2;
              if (mypass != pass)
                continue;
              const int64_t nb = H.nbr[cell * 2 * dim + 2 * d + 1];
              const int     t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
              for (int fq = 0; fq < mf; ++fq)
                {
                  const int a = fq % m, b = dim == 3 ? fq / m : 0;
                  double    xm[3] = {0, 0, 0}, xp[3] = {0, 0, 0};
                  xm[d]  = 1;
                  xp[d]  = 0;
                  xm[t1] = xp[t1] = (double)gx[a];
                  double w = (double)gw[a];
                  if (dim == 3)
                    {
                      xm[t2] = xp[t2] = (double)gx[b];
                      w *= (double)gw[b];
                    }
                  double x[3], J[3][3];
                  map_point(H, cell, xm, x, J);
                  double nv[3] = {0, 0, 0};
                  if (dim == 3)
                    {
                      const int p1 = (d + 1) % 3, p2 = (d + 2) % 3;
                      nv[0] = J[1][p1] * J[2][p2] - J[2][p1] * J[1][p2];
                      nv[1] = J[2][p1] * J[0][p2] - J[0][p1] * J[2][p2];
                      nv[2] = J[0][p1] * J[1][p2] - J[1][p1] * J[0][p2];
                    }
                  else
                    {
                      const int t = 1 - d;
                      nv[0]       = d == 0 ? J[1][t] : -J[1][t];
                      nv[1]       = d == 0 ? -J[0][t] : J[0][t];
                    }
                  double len = 0;
                  for (int e = 0; e < dim; ++e)
                    len += nv[e] * nv[e];
                  len = std::sqrt(len);
                  double nrm[3];
                  for (int e = 0; e < dim; ++e)
                    nrm[e] = nv[e] / len;
                  const double JxW = len * w;
                  double       um[3], up[3], gr[3];
                  for (int c = 0; c < nc; ++c)
                    {
                      point_eval(H, src + (cell * nc + c) * nd, xm, um[c], gr);
                      point_eval(H, src + (nb * nc + c) * nd, xp, up[c], gr);
                    }
                  double umn = 0, upn = 0;
                  for (int e = 0; e < dim; ++e)
                    {
                      umn += um[e] * nrm[e];
                      upn += up[e] * nrm[e];
                    }
                  const double lam =
                    H.P.flux_alpha * std::max(std::fabs(umn), std::fabs(upn));
                  for (int c = 0; c < nc; ++c)
                    {
                      const double flux =
                        0.5 * ((umn * um[c] + upn * up[c]) +
                               lam * (um[c] - up[c]));
                      double Jd[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                      point_integrate(H, xm, Jd, nrm, -flux * JxW, 0.,
                                      dst + (cell * nc + c) * nd);
                      point_integrate(H, xp, Jd, nrm, flux * JxW, 0.,
                                      dst + (nb * nc + c) * nd);
                    }
                }
            }
        }
  }

  // SIPG face terms of advection_diffusion_solver.cc:569-613 (interior
  // faces, the minus side's normal on both sides)
  void
  sipg_face_terms(const or_handle &H, const LoopSpec &L, const double *src,
                  double *dst)
  {
    const int    dim = H.dim, n = H.n, nd = H.nd;
    const int    nf  = dim == 3 ? n * n : n;
    const double nu  = H.P.viscosity;
    const double sg  = L.implicit ? 1. : -1.;
    for (int d = 0; d < dim; ++d)
      for (int pass = 0; pass < 3; ++pass)
        {
#pragma omp parallel for schedule(static)
          for (int64_t cell = 0; cell < H.n_cells; ++cell)
            {
              const int id = H.ijk[3 * cell + d];
              const int mypass =
                (H.N[d] % 2 == 1 && id == H.N[d] - 1) ? 2 : id % 2;
              if (mypass != pass)
                continue;
              const int64_t nb = H.nbr[cell * 2 * dim + 2 * d + 1];
              if (nb < 0)
                continue;
              const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
              for (int fq = 0; fq < nf; ++fq)
                {
                  int    tq[2] = {fq % n, dim == 3 ? fq / n : 0};
                  double x[3], normal[3], JxW;
                  face_geometry(H, cell, d, 1, tq, x, normal, JxW);
                  double xm[3] = {0, 0, 0}, xp[3] = {0, 0, 0};
                  xm[d] = 1;
                  xp[d] = 0;
                  xm[t1] = xp[t1] = H.xq[tq[0]];
                  if (dim == 3)
                    xm[t2] = xp[t2] = H.xq[tq[1]];
                  double Jm[3][3], Jp[3][3], Jim[3][3], Jip[3][3], xx[3];
                  map_point(H, cell, xm, xx, Jm);
                  map_point(H, nb, xp, xx, Jp);
                  inverse_jacobian(dim, Jm, Jim);
                  inverse_jacobian(dim, Jp, Jip);
                  double um, up, gm[3], gp[3];
                  point_eval(H, src + cell * nd, xm, um, gm);
                  point_eval(H, src + nb * nd, xp, up, gp);
                  double dnm = 0, dnp = 0, ihm = 0, ihp = 0;
                  for (int e = 0; e < dim; ++e)
                    {
                      double gme = 0, gpe = 0;
                      for (int g = 0; g < dim; ++g)
                        {
                          gme += Jim[g][e] * gm[g];
                          gpe += Jip[g][e] * gp[g];
                        }
                      dnm += normal[e] * gme;
                      dnp += normal[e] * gpe;
                      ihm += Jim[d][e] * normal[e];
                      ihp += Jip[d][e] * normal[e];
                    }
                  const double sigma =
                    H.P.penalty_factor * 0.5 * (std::fabs(ihm) + std::fabs(ihp));
                  const double jump    = (um - up) * (-0.5 * nu);
                  const double avg     = (dnm + dnp) * (-0.5 * nu);
                  const double viscous = jump * sigma - avg;
                  point_integrate(H, xm, Jim, normal, -viscous * JxW * sg,
                                  jump * JxW * sg, dst + cell * nd);
                  point_integrate(H, xp, Jip, normal, viscous * JxW * sg,
                                  jump * JxW * sg, dst + nb * nd);
                }
            }
        }
  }

  void
  run_loop(const or_handle &H, const LoopSpec &L, const double *src,
           int64_t block, double *dst, double t_bc)
  {
    std::fill(dst, dst + block * L.n_stages, 0.);
#pragma omp parallel
    {
      CellEval E;
#pragma omp for schedule(static)
      for (int64_t cell = 0; cell < H.n_cells; ++cell)
        cell_term(H, L, cell, src, block, dst, E);
    }
    face_terms(H, L, src, block, dst, t_bc);
    if (H.P.operator_kind == 1)
      sipg_face_terms(H, L, src, dst);
  }

  void
  inverse_mass_cells(const or_handle &H, const double *src, double *dst)
  {
    const int n = H.n, nd = H.nd, dim = H.dim;
#pragma omp parallel
    {
      std::vector<double> a(nd), b(nd), t1, t2;
#pragma omp for schedule(static)
      for (int64_t cell = 0; cell < H.n_cells; ++cell)
        {
          // S^-T: (S^-1)^T applied as the transpose of Sinv, with S the
          // nodal basis at the mass matrix's quadrature points
          const double *Mi[3] = {H.Sminv.data(), H.Sminv.data(),
                                 H.Sminv.data()};
          tensor_apply(dim, Mi, n, n, true, src + cell * nd, a.data(), t1, t2);
          for (int q = 0; q < nd; ++q)
            {
              int qi[3];
              q_index(H, q, qi);
              double xi[3] = {H.xqm[qi[0]], H.xqm[qi[1]],
                              dim == 3 ? H.xqm[qi[2]] : 0};
              double x[3], J[3][3], Ji[3][3];
              map_point(H, cell, xi, x, J);
              const double det = inverse_jacobian(dim, J, Ji);
              double       w   = H.wqm[qi[0]] * H.wqm[qi[1]];
              if (dim == 3)
                w *= H.wqm[qi[2]];
              a[q] /= det * w;
            }
          tensor_apply(dim, Mi, n, n, false, a.data(), b.data(), t1, t2);
          std::copy(b.begin(), b.end(), dst + cell * nd);
        }
    }
  }

  std::vector<double>
  lsrk_b(int scheme, std::vector<double> &ai)
  {
    // advection_solver.cc:828-897
    std::vector<double> bi;
    switch (scheme)
      {
        case 0:
          bi = {0.245170287303492, 0.184896052186740, 0.569933660509768};
          ai = {0.755726351946097, 0.386954477304099};
          break;
        case 1:
          bi = {1153189308089. / 22510343858157., 1772645290293. / 4653164025191.,
                -1672844663538. / 4480602732383., 2114624349019. / 3568978502595.,
                5198255086312. / 14908931495163.};
          ai = {970286171893. / 4311952581923., 6584761158862. / 12103376702013.,
                2251764453980. / 15575788980749., 26877169314380. / 34165994151039.};
          break;
        case 2:
          bi = {0.0941840925477795334, 0.149683694803496998, 0.285204742060440058,
                -0.122201846148053668, 0.0605151571191401122, 0.345986987898399296,
                0.186627171718797670};
          ai = {0.241566650129646868 + bi[0], 0.0423866513027719953 + bi[1],
                0.215602732678803776 + bi[2], 0.232328007537583987 + bi[3],
                0.256223412574146438 + bi[4], 0.0978694102142697230 + bi[5]};
          break;
        default:
          bi = {2274579626619. / 23610510767302., 693987741272. / 12394497460941.,
                -347131529483. / 15096185902911., 1144057200723. / 32081666971178.,
                1562491064753. / 11797114684756., 13113619727965. / 44346030145118.,
                393957816125. / 7825732611452., 720647959663. / 6565743875477.,
                3559252274877. / 14424734981077.};
          ai = {1107026461565. / 5417078080134., 38141181049399. / 41724347789894.,
                493273079041. / 11940823631197., 1851571280403. / 6147804934346.,
                11782306865191. / 62590030070788., 9452544825720. / 13648368537481.,
                4435885630781. / 26285702406235., 2357909744247. / 11371140753790.};
      }
    return bi;
  }

  // Radau IIA from its definition: c_i are the zeros of
  // d^{s-1}/dx^{s-1} [x^{s-1} (x-1)^s] (c_s = 1), A_ij = int_0^{c_i} l_j.
  void
  radau_iia(int s, std::vector<double> &A, std::vector<double> &b,
            std::vector<double> &c)
  {
    // polynomial q(x) = x^{s-1}(x-1)^s, coefficients ascending
    std::vector<ld> q(2 * s, 0.);
    q[s - 1] = 1;
    for (int k = 0; k < s; ++k)
      {
        std::vector<ld> r(q.size(), 0.);
        for (size_t i = 0; i + 1 < q.size(); ++i)
          {
            r[i + 1] += q[i];
            r[i] -= q[i];
          }
        q = r;
      }
    for (int k = 0; k < s - 1; ++k)
      {
        std::vector<ld> r(q.size(), 0.);
        for (size_t i = 1; i < q.size(); ++i)
          r[i - 1] = q[i] * (ld)i;
        q = r;
      }
    // the s-1 interior roots by sampling + bisection, then c_s = 1
    std::vector<ld> roots;
    auto            ev = [&](ld x) {
      ld v = 0;
      for (int i = (int)q.size() - 1; i >= 0; --i)
        v = v * x + q[i];
      return v;
    };
    const int M = 20000;
    for (int i = 0; i < M && (int)roots.size() < s - 1; ++i)
      {
        ld a = (ld)i / M * (1 - 1e-9L), bb = (ld)(i + 1) / M * (1 - 1e-9L);
        ld fa = ev(a), fb = ev(bb);
        if ((fa < 0) != (fb < 0))
          {
            for (int it = 0; it < 200; ++it)
              {
                ld m = (a + bb) / 2;
                if ((ev(m) < 0) == (fa < 0))
                  a = m;
                else
                  bb = m;
              }
            roots.push_back((a + bb) / 2);
          }
      }
    roots.push_back(1);
    c.assign(s, 0.);
    A.assign(s * s, 0.);
    b.assign(s, 0.);
    for (int i = 0; i < s; ++i)
      c[i] = (double)roots[i];
    // integrate Lagrange polynomials with a 12-point Gauss rule
    std::vector<ld> gx, gw;
    gauss(12, gx, gw);
    for (int i = 0; i < s; ++i)
      for (int j = 0; j < s; ++j)
        {
          ld sum = 0;
          for (size_t g = 0; g < gx.size(); ++g)
            {
              const ld x = gx[g] * roots[i];
              sum += gw[g] * roots[i] * lagr(roots, j, x);
            }
          A[i * s + j] = (double)sum;
        }
    for (int j = 0; j < s; ++j)
      b[j] = A[(s - 1) * s + j];
  }
} // namespace

extern "C" {

void
or_set_threads(int n)
{
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

void
or_basis_1d(int n, double *gp, double *gw, double *gl)
{
  std::vector<ld> x, w;
  gauss(n, x, w);
  const auto l = lobatto(n);
  for (int i = 0; i < n; ++i)
    {
      gp[i] = (double)x[i];
      gw[i] = (double)w[i];
      gl[i] = (double)l[i];
    }
}

or_handle *
or_create(const or_problem *p)
{
  auto *H = new or_handle();
  H->P    = *p;
  H->dim  = p->dim;
  H->n    = p->degree + 1;
  H->nd   = p->dim == 3 ? H->n * H->n * H->n : H->n * H->n;
  build_mesh(*H);
  const int       n = H->n;
  std::vector<ld> x, w;
  gauss(n, x, w);
  H->xgl = lobatto(n);
  // Gauss-Lobatto weights: the integrals of the nodal Lagrange polynomials
  // (degree n-1), by the n-point Gauss rule
  std::vector<ld> wgl(n, 0);
  for (int i = 0; i < n; ++i)
    for (int q = 0; q < n; ++q)
      wgl[i] += w[q] * lagr(H->xgl, i, x[q]);
  const bool gl_op = p->use_gl_quad != 0;
  const bool gl_m  = gl_op || p->use_gl_quad_mass != 0;
  if (gl_op)
    {
      x = H->xgl;
      w = wgl;
    }
  H->xq.resize(n);
  H->wq.resize(n);
  for (int i = 0; i < n; ++i)
    {
      H->xq[i] = (double)x[i];
      H->wq[i] = (double)w[i];
    }
  H->S.resize(n * n);
  H->G.resize(n * n);
  H->phi0.resize(n);
  H->phi1.resize(n);
  for (int q = 0; q < n; ++q)
    for (int i = 0; i < n; ++i)
      {
        H->S[q * n + i] = (double)lagr(H->xgl, i, x[q]);
        H->G[q * n + i] = (double)lagr_der(H->xgl, i, x[q]);
      }
  for (int i = 0; i < n; ++i)
    {
      H->phi0[i] = (double)lagr(H->xgl, i, 0);
      H->phi1[i] = (double)lagr(H->xgl, i, 1);
    }
  H->Sinv = invert(H->S, n);
  if (gl_m && !gl_op)
    {
      H->xqm.resize(n);
      H->wqm.resize(n);
      H->Sm.resize(n * n);
      for (int q = 0; q < n; ++q)
        {
          H->xqm[q] = (double)H->xgl[q];
          H->wqm[q] = (double)wgl[q];
          for (int i = 0; i < n; ++i)
            H->Sm[q * n + i] = (double)lagr(H->xgl, i, H->xgl[q]);
        }
      H->Sminv = invert(H->Sm, n);
    }
  else
    {
      H->xqm   = H->xq;
      H->wqm   = H->wq;
      H->Sm    = H->S;
      H->Sminv = H->Sinv;
    }
  return H;
}

void
or_destroy(or_handle *h)
{
  delete h;
}

int64_t
or_n_cells(const or_handle *h)
{
  return h->n_cells;
}
int64_t
or_n_dofs(const or_handle *h)
{
  return h->n_cells * h->nd;
}

void
or_cell_coordinates(const or_handle *h, int32_t *ijk)
{
  for (size_t i = 0; i < h->ijk.size(); ++i)
    ijk[i] = h->ijk[i];
}
void
or_neighbors(const or_handle *h, int32_t *nb)
{
  for (size_t i = 0; i < h->nbr.size(); ++i)
    nb[i] = h->nbr[i];
}
void
or_cell_vertices(const or_handle *h, double *v)
{
  std::memcpy(v, h->corners.data(), h->corners.size() * sizeof(double));
}
double
or_min_vertex_distance(const or_handle *h)
{
  return h->min_dist;
}

void
or_loop(or_handle *h, const double *src, double *dst, double t)
{
  LoopSpec L;
  L.implicit       = false;
  L.inv_dt         = 0;
  L.n_stages       = 1;
  L.time_factor[0] = h->time_factor(t);
  run_loop(*h, L, src, h->n_cells * h->nd, dst, t);
}

void
or_inverse_mass(or_handle *h, const double *src, double *dst)
{
  inverse_mass_cells(*h, src, dst);
}

void
or_mass(or_handle *h, const double *src, double *dst)
{
  // the mass matrix of the mass quadrature: S^T diag(JxW) S
  const int           n = h->n, nd = h->nd, dim = h->dim;
  std::vector<double> a(nd), t1, t2;
  const double       *Ms[3] = {h->Sm.data(), h->Sm.data(), h->Sm.data()};
  for (int64_t cell = 0; cell < h->n_cells; ++cell)
    {
      tensor_apply(dim, Ms, n, n, false, src + cell * nd, a.data(), t1, t2);
      for (int q = 0; q < nd; ++q)
        {
          int qi[3];
          q_index(*h, q, qi);
          double xi[3] = {h->xqm[qi[0]], h->xqm[qi[1]],
                          dim == 3 ? h->xqm[qi[2]] : 0};
          double x[3], J[3][3], Ji[3][3];
          map_point(*h, cell, xi, x, J);
          const double det = inverse_jacobian(dim, J, Ji);
          double       w   = h->wqm[qi[0]] * h->wqm[qi[1]] *
                     (dim == 3 ? h->wqm[qi[2]] : 1);
          a[q] *= det * w;
        }
      tensor_apply(dim, Ms, n, n, true, a.data(), dst + cell * nd, t1, t2);
    }
}

void
or_apply(or_handle *h, const double *src, double *dst, double t)
{
  or_loop(h, src, dst, t);
  std::vector<double> tmp(dst, dst + h->n_cells * h->nd);
  inverse_mass_cells(*h, tmp.data(), dst);
}

void
or_perform_stage(or_handle *h, double t, double b_dt, double a_dt,
                 const double *current_ri, double *vec_ki, double *solution,
                 double *next_ri)
{
  const int64_t N = h->n_cells * h->nd;
  or_loop(h, current_ri, vec_ki, t);
  std::vector<double> k(N);
  inverse_mass_cells(*h, vec_ki, k.data());
  for (int64_t i = 0; i < N; ++i)
    {
      const double sol = solution[i];
      solution[i]      = sol + b_dt * k[i];
      if (a_dt != 0.)
        next_ri[i] = sol + a_dt * k[i];
      else
        next_ri[i] = k[i];
    }
}

void
or_lsrk_time_step(or_handle *h, int scheme, double t, double dt,
                  double *solution, double *vec_ri, double *vec_ki)
{
  std::vector<double> ai;
  std::vector<double> bi = lsrk_b(scheme, ai);
  or_perform_stage(h, t, bi[0] * dt, ai[0] * dt, solution, vec_ri, solution,
                   vec_ri);
  double sum_previous_bi = 0;
  for (size_t stage = 1; stage < bi.size(); ++stage)
    {
      const double c_i = sum_previous_bi + ai[stage - 1];
      or_perform_stage(h, t + c_i * dt, bi[stage] * dt,
                       stage == bi.size() - 1 ? 0 : ai[stage] * dt, vec_ri,
                       vec_ki, solution, vec_ri);
      sum_previous_bi += bi[stage - 1];
    }
}

void
or_project(or_handle *h, double *dst, double t)
{
  const int n = h->n, nd = h->nd, dim = h->dim, nc = h->P.n_components;
  std::vector<double> v(nd), t1, t2;
  for (int64_t cell = 0; cell < h->n_cells; ++cell)
    for (int c = 0; c < nc; ++c)
      {
        for (int q = 0; q < nd; ++q)
          {
            int qi[3];
            q_index(*h, q, qi);
            double xi[3] = {h->xq[qi[0]], h->xq[qi[1]], dim == 3 ? h->xq[qi[2]] : 0};
            double x[3], J[3][3];
            map_point(*h, cell, xi, x, J);
            v[q] = solution(h->P, x, t, c);
          }
        const double *Mi[3] = {h->Sinv.data(), h->Sinv.data(), h->Sinv.data()};
        tensor_apply(dim, Mi, n, n, false, v.data(), dst + (cell * nc + c) * nd,
                     t1, t2);
      }
}

void
or_mass_energy(or_handle *h, const double *vec, double out[3])
{
  const int nd = h->nd, dim = h->dim;
  double    m = 0, e = 0, h1 = 0;
  for (int64_t cell = 0; cell < h->n_cells; ++cell)
    {
      CellEval E;
      evaluate(*h, vec + cell * nd, E, true);
      for (int q = 0; q < nd; ++q)
        {
          int qi[3];
          q_index(*h, q, qi);
          double xi[3] = {h->xq[qi[0]], h->xq[qi[1]], dim == 3 ? h->xq[qi[2]] : 0};
          double x[3], J[3][3], Ji[3][3];
          map_point(*h, cell, xi, x, J);
          const double det = inverse_jacobian(dim, J, Ji);
          const double w =
            h->wq[qi[0]] * h->wq[qi[1]] * (dim == 3 ? h->wq[qi[2]] : 1);
          const double JxW = det * w;
          double       g2  = 0;
          for (int a = 0; a < dim; ++a)
            {
              double g = 0;
              for (int d = 0; d < dim; ++d)
                g += Ji[d][a] * E.grads[d][q];
              g2 += g * g;
            }
          m += E.vals[q] * JxW;
          e += E.vals[q] * E.vals[q] * JxW;
          h1 += g2 * JxW;
        }
    }
  out[0] = m;
  out[1] = e;
  out[2] = h1;
}

void
or_l2_error(or_handle *h, const double *vec, double t, double out[2])
{
  // brute force: evaluate the nodal expansion directly at the points of
  // QGauss(k+2) (error) and QGauss(k+1) (norm)
  const int dim = h->dim, n = h->n, nd = h->nd;
  double    err = 0, nrm = 0;
  for (int rule = 0; rule < 2; ++rule)
    {
      const int       m = rule == 0 ? n + 1 : n;
      std::vector<ld> x, w;
      gauss(m, x, w);
      std::vector<double> phi(m * n);
      for (int q = 0; q < m; ++q)
        for (int i = 0; i < n; ++i)
          phi[q * n + i] = (double)lagr(h->xgl, i, x[q]);
      const int npts = dim == 3 ? m * m * m : m * m;
      for (int64_t cell = 0; cell < h->n_cells; ++cell)
        for (int q = 0; q < npts; ++q)
          {
            int qi[3] = {q % m, (q / m) % m, dim == 3 ? q / (m * m) : 0};
            double xi[3] = {(double)x[qi[0]], (double)x[qi[1]],
                            dim == 3 ? (double)x[qi[2]] : 0};
            double xx[3], J[3][3], Ji[3][3];
            map_point(*h, cell, xi, xx, J);
            const double det = inverse_jacobian(dim, J, Ji);
            double       wq  = (double)(w[qi[0]] * w[qi[1]]);
            if (dim == 3)
              wq *= (double)w[qi[2]];
            const double g = solution(h->P, xx, t);
            if (rule == 0)
              {
                double uh = 0;
                for (int i = 0; i < nd; ++i)
                  {
                    int ii[3] = {i % n, (i / n) % n, dim == 3 ? i / (n * n) : 0};
                    double f = phi[qi[0] * n + ii[0]] * phi[qi[1] * n + ii[1]];
                    if (dim == 3)
                      f *= phi[qi[2] * n + ii[2]];
                    uh += f * vec[cell * nd + i];
                  }
                err += (uh - g) * (uh - g) * det * wq;
              }
            else
              nrm += g * g * det * wq;
          }
    }
  out[0] = std::sqrt(err);
  out[1] = std::sqrt(nrm);
}

void
or_vmult(or_handle *h, double t, double dt, const double *src, double *dst)
{
  // advection_solver_implicit_real.cc:876-891 with factor_time at t + dt
  LoopSpec L;
  L.implicit       = true;
  L.inv_dt         = dt == 0 ? 0. : 1. / dt;
  L.n_stages       = 1;
  L.time_factor[0] = h->time_factor(t + dt);
  run_loop(*h, L, src, h->n_cells * h->nd, dst, t + dt);
}

void
or_irk_coefficients(int s, double *A, double *b, double *c)
{
  std::vector<double> a, bb, cc;
  radau_iia(s, a, bb, cc);
  for (int i = 0; i < s * s; ++i)
    A[i] = a[i];
  for (int i = 0; i < s; ++i)
    {
      b[i] = bb[i];
      c[i] = cc[i];
    }
}

void
or_irk_compute_rhs(or_handle *h, double t, double dt, const double *src,
                   double *dst)
{
  // dst_c = -L(t + c_c dt) src / dt (advection_solver_irk_real.cc:1266-1282)
  const int           s = h->P.irk_stages;
  std::vector<double> A, b, c;
  radau_iia(s, A, b, c);
  const int64_t N = h->n_cells * h->nd;
  for (int st = 0; st < s; ++st)
    {
      LoopSpec L;
      L.implicit       = true;
      L.inv_dt         = 0;
      L.n_stages       = 1;
      L.time_factor[0] = h->time_factor(t + c[st] * dt);
      run_loop(*h, L, src, N, dst + st * N, t + c[st] * dt);
      for (int64_t i = 0; i < N; ++i)
        dst[st * N + i] *= -1. / dt;
    }
}

void
or_irk_vmult(or_handle *h, double t, double dt, const double *src,
             double *dst)
{
  // (A^-1 (x) M)/dt Z + diag(L(t + c dt)) Z (advection_solver_irk_real.cc:1561-1703)
  const int           s = h->P.irk_stages;
  std::vector<double> A, b, c;
  radau_iia(s, A, b, c);
  const auto invA = invert(A, s);
  LoopSpec   L;
  L.implicit = true;
  L.inv_dt   = 1. / dt;
  L.n_stages = s;
  L.mass.resize(s * s);
  for (int i = 0; i < s * s; ++i)
    L.mass[i] = invA[i] / dt;
  for (int st = 0; st < s; ++st)
    L.time_factor[st] = h->time_factor(t + c[st] * dt);
  run_loop(*h, L, src, h->n_cells * h->nd, dst, t);
}

void
or_irk_update_solution(or_handle *h, double dt, double *solution,
                       const double *stage)
{
  const int           s = h->P.irk_stages;
  std::vector<double> A, b, c;
  radau_iia(s, A, b, c);
  const auto    invA = invert(A, s);
  const int64_t N    = h->n_cells * h->nd;
  for (int64_t i = 0; i < N; ++i)
    {
      double my = 0;
      for (int cc = 0; cc < s; ++cc)
        {
          double sum = 0;
          for (int bb = 0; bb < s; ++bb)
            sum += stage[bb * N + i] * invA[cc * s + bb];
          my += b[cc] * sum;
        }
      solution[i] += dt * my;
    }
}

} // extern "C"

extern "C" {

void
or_burgers_apply(or_handle *h, const double *src, double *dst)
{
  // BurgersOperation::apply: loop then inverse mass per component
  // (burgers_solver.cc:386-410)
  const int64_t nc = h->P.n_components, nd = h->nd;
  std::vector<double> tmp(h->n_cells * nc * nd);
  burgers_loop(*h, src, tmp.data());
  std::vector<double> a(h->n_cells * nd), b(h->n_cells * nd);
  for (int64_t c = 0; c < nc; ++c)
    {
      for (int64_t cell = 0; cell < h->n_cells; ++cell)
        std::copy(tmp.begin() + (cell * nc + c) * nd,
                  tmp.begin() + (cell * nc + c + 1) * nd, a.begin() + cell * nd);
      inverse_mass_cells(*h, a.data(), b.data());
      for (int64_t cell = 0; cell < h->n_cells; ++cell)
        std::copy(b.begin() + cell * nd, b.begin() + (cell + 1) * nd,
                  dst + (cell * nc + c) * nd);
    }
}

} // extern "C"
