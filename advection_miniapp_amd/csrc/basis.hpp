// One-dimensional polynomial data for the DG operator: Gauss and
// Gauss-Lobatto point sets on [0,1], Lagrange bases on them, and the small
// n x n matrices the sum-factorized kernels contract with.
//
// The element is FE_DGQ(k) of deal.II: Lagrange polynomials on the k+1
// Gauss-Lobatto nodes, lexicographic numbering (advection_solver.cc:961,
// `FE_DGQ<dim> fe`), integrated with QGauss(k+1), or with
// QGaussLobatto(k+1) under use_gl_quad (advection_solver.cc:110-116,
// :440-445).
#pragma once

#include <cmath>
#include <stdexcept>
#include <vector>

namespace dgops
{
  constexpr int max_n = 8; // degree <= 7

  using real_ext = long double;

  // Legendre polynomial P_m and its first two derivatives at x in [-1,1].
  inline void
  legendre(int m, real_ext x, real_ext &p, real_ext &dp, real_ext &ddp)
  {
    real_ext p0 = 1, p1 = x;
    if (m == 0)
      {
        p = 1;
        dp = 0;
        ddp = 0;
        return;
      }
    for (int k = 2; k <= m; ++k)
      {
        const real_ext pk = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = pk;
      }
    p = p1;
    // (1-x^2) P' = m (P_{m-1} - x P_m)
    dp = m * (p0 - x * p1) / (1 - x * x);
    // (1-x^2) P'' = 2x P' - m(m+1) P
    ddp = (2 * x * dp - m * (m + 1) * p) / (1 - x * x);
  }

  // Gauss-Legendre points and weights on [0,1], ascending.
  inline void
  gauss_points(int n, std::vector<double> &pts, std::vector<double> &wts)
  {
    pts.assign(n, 0.);
    wts.assign(n, 0.);
    for (int i = 0; i < n; ++i)
      {
        // root i of P_n counted from the left
        real_ext x = -std::cos(M_PI * (i + 0.75L) / (n + 0.5L));
        for (int it = 0; it < 100; ++it)
          {
            real_ext p, dp, ddp;
            legendre(n, x, p, dp, ddp);
            const real_ext dx = p / dp;
            x -= dx;
            if (std::fabs((double)dx) < 1e-19)
              break;
          }
        real_ext p, dp, ddp;
        legendre(n, x, p, dp, ddp);
        pts[i] = (double)((1 + x) / 2);
        wts[i] = (double)(1 / ((1 - x * x) * dp * dp));
      }
  }

  // Gauss-Lobatto points on [0,1]: endpoints plus the roots of P'_{n-1}.
  inline std::vector<real_ext>
  gauss_lobatto_points_ext(int n)
  {
    std::vector<real_ext> pts(n);
    if (n == 1)
      {
        pts[0] = 0.5L;
        return pts;
      }
    pts[0] = 0;
    pts[n - 1] = 1;
    const int m = n - 1;
    for (int i = 1; i < n - 1; ++i)
      {
        real_ext x = -std::cos(M_PI * i / (real_ext)m);
        for (int it = 0; it < 100; ++it)
          {
            real_ext p, dp, ddp;
            legendre(m, x, p, dp, ddp);
            const real_ext dx = dp / ddp;
            x -= dx;
            if (std::fabs((double)dx) < 1e-19)
              break;
          }
        pts[i] = (1 + x) / 2;
      }
    return pts;
  }

  inline std::vector<real_ext>
  gauss_points_ext(int n)
  {
    std::vector<real_ext> pts(n);
    for (int i = 0; i < n; ++i)
      {
        real_ext x = -std::cos(M_PI * (i + 0.75L) / (n + 0.5L));
        for (int it = 0; it < 100; ++it)
          {
            real_ext p, dp, ddp;
            legendre(n, x, p, dp, ddp);
            const real_ext dx = p / dp;
            x -= dx;
            if (std::fabs((double)dx) < 1e-19)
              break;
          }
        pts[i] = (1 + x) / 2;
      }
    return pts;
  }

  // Lagrange polynomial i on `nodes`, value and derivative at x.
  inline void
  lagrange(const std::vector<real_ext> &nodes,
           int                          i,
           real_ext                     x,
           real_ext                    &val,
           real_ext                    &der)
  {
    const int n = (int)nodes.size();
    val = 1;
    for (int j = 0; j < n; ++j)
      if (j != i)
        val *= (x - nodes[j]) / (nodes[i] - nodes[j]);
    der = 0;
    for (int k = 0; k < n; ++k)
      {
        if (k == i)
          continue;
        real_ext prod = 1 / (nodes[i] - nodes[k]);
        for (int j = 0; j < n; ++j)
          if (j != i && j != k)
            prod *= (x - nodes[j]) / (nodes[i] - nodes[j]);
        der += prod;
      }
  }

  // Gauss-Lobatto weights on [0,1]: 1 / (n (n-1) P_{n-1}(x_i)^2) at the
  // points of gauss_lobatto_points_ext (x_i mapped to [-1,1])
  inline std::vector<real_ext>
  gauss_lobatto_weights_ext(int n)
  {
    std::vector<real_ext> w(n, (real_ext)1);
    if (n == 1)
      return w;
    const auto x = gauss_lobatto_points_ext(n);
    for (int i = 0; i < n; ++i)
      {
        real_ext p, dp, ddp;
        legendre(n - 1, 2 * x[i] - 1, p, dp, ddp);
        w[i] = 1 / ((real_ext)n * (n - 1) * p * p);
      }
    return w;
  }

  // All 1D data one degree needs. Matrices are row-major [row][col] with
  // the row index first, padded to max_n. The quadrature ("Gauss") points
  // and weights are Gauss-Legendre, or Gauss-Lobatto for the
  // use_gl_quad switch: then they coincide with the nodes, S and S^-1 are
  // the identity and l0, l1 pick the end nodes.
  struct Basis1D
  {
    int    n;
    bool   gl_quadrature;
    double gauss_pts[max_n];
    double gauss_wts[max_n];
    double gl_pts[max_n];
    // S[q][i]    = phi_i(xg_q): nodal (Gauss-Lobatto) basis at Gauss points
    double S[max_n][max_n];
    // Sinv[i][q] = inverse of S = Gauss-Lagrange ell_q at node i
    double Sinv[max_n][max_n];
    // D[q][p]    = ell_p'(xg_q): collocation derivative at Gauss points
    double D[max_n][max_n];
    // G[q][i]    = phi_i'(xg_q): derivative of the nodal basis at Gauss pts
    double G[max_n][max_n];
    // Gauss-Lagrange values and derivatives at the two cell faces
    double l0[max_n], l1[max_n], dl0[max_n], dl1[max_n];
    // derivatives of the nodal basis at the two faces
    double dphi0[max_n], dphi1[max_n];
  };

  inline Basis1D
  make_basis(int n, bool gl_quadrature = false)
  {
    if (n < 1 || n > max_n)
      throw std::runtime_error("unsupported number of 1D points");
    if (gl_quadrature && n < 2)
      throw std::runtime_error("Gauss-Lobatto quadrature needs degree >= 1");
    Basis1D b = {};
    b.n             = n;
    b.gl_quadrature = gl_quadrature;
    std::vector<double> gp, gw;
    gauss_points(n, gp, gw);
    const auto xl = gauss_lobatto_points_ext(n);
    const auto xg = gl_quadrature ? xl : gauss_points_ext(n);
    if (gl_quadrature)
      {
        const auto wl = gauss_lobatto_weights_ext(n);
        for (int q = 0; q < n; ++q)
          {
            gp[q] = (double)xl[q];
            gw[q] = (double)wl[q];
          }
      }
    for (int q = 0; q < n; ++q)
      {
        b.gauss_pts[q] = gp[q];
        b.gauss_wts[q] = gw[q];
        b.gl_pts[q]    = (double)xl[q];
      }
    for (int q = 0; q < n; ++q)
      for (int i = 0; i < n; ++i)
        {
          real_ext v, d;
          lagrange(xl, i, xg[q], v, d);
          b.S[q][i] = (double)v;
          b.G[q][i] = (double)d;
          lagrange(xg, i, xg[q], v, d);
          b.D[q][i] = (double)d;
          lagrange(xg, q, xl[i], v, d);
          b.Sinv[i][q] = (double)v;
        }
    for (int q = 0; q < n; ++q)
      {
        real_ext v, d;
        lagrange(xg, q, 0.0L, v, d);
        b.l0[q]  = (double)v;
        b.dl0[q] = (double)d;
        lagrange(xg, q, 1.0L, v, d);
        b.l1[q]  = (double)v;
        b.dl1[q] = (double)d;
        lagrange(xl, q, 0.0L, v, d);
        b.dphi0[q] = (double)d;
        lagrange(xl, q, 1.0L, v, d);
        b.dphi1[q] = (double)d;
      }
    return b;
  }
} // namespace dgops
