// Geometry and velocity along the 1D lines the operator kernel owns.
//
// On a multilinear (straight-sided, perturbed) cell the position x and the
// Jacobian columns are linear along every reference coordinate line, so a
// thread evaluates the map only at the two ends of its line and
// interpolates linearly to the Gauss points. The analytic velocity fields
// are products of sin/cos(pi c x_e); along a line pi c x_e = theta0 +
// t dtheta, and for small dtheta the values at the points follow from
// sincospi at the start and a short Taylor series in t dtheta (error below
// 1e-17 for |t dtheta| < 1/4), otherwise sincospi is called per point.
#pragma once

#include "analytic.hpp"
#include "dev_basis.hpp"
#include "dg_common.hpp"

namespace dgops
{
  // Multilinear map of a cell from its 2^dim corners X[v*dim+e]:
  // position x and Jacobian J[e][d] = dx_e / dxi_d at reference point xi.
  template <int dim>
  __device__ __forceinline__ void
  multilinear_map(const double *X,
                  const double (&xi)[3],
                  double (&x)[3],
                  double (&J)[3][3])
  {
    if constexpr (dim == 3)
      {
        const double a = xi[0], b = xi[1], c = xi[2];
        const double ma = 1 - a, mb = 1 - b, mc = 1 - c;
        for (int e = 0; e < 3; ++e)
          {
            const double x0 = X[0 * 3 + e], x1 = X[1 * 3 + e],
                         x2 = X[2 * 3 + e], x3 = X[3 * 3 + e],
                         x4 = X[4 * 3 + e], x5 = X[5 * 3 + e],
                         x6 = X[6 * 3 + e], x7 = X[7 * 3 + e];
            const double e01 = ma * x0 + a * x1, e23 = ma * x2 + a * x3,
                         e45 = ma * x4 + a * x5, e67 = ma * x6 + a * x7;
            const double f0 = mb * e01 + b * e23, f1 = mb * e45 + b * e67;
            x[e] = mc * f0 + c * f1;
            J[e][0] = mc * (mb * (x1 - x0) + b * (x3 - x2)) +
                      c * (mb * (x5 - x4) + b * (x7 - x6));
            J[e][1] = mc * (e23 - e01) + c * (e67 - e45);
            J[e][2] = f1 - f0;
          }
      }
    else
      {
        const double a = xi[0], b = xi[1];
        const double ma = 1 - a, mb = 1 - b;
        for (int e = 0; e < 2; ++e)
          {
            const double x0 = X[0 * 2 + e], x1 = X[1 * 2 + e],
                         x2 = X[2 * 2 + e], x3 = X[3 * 2 + e];
            const double e01 = ma * x0 + a * x1, e23 = ma * x2 + a * x3;
            x[e]    = mb * e01 + b * e23;
            J[e][0] = mb * (x1 - x0) + b * (x3 - x2);
            J[e][1] = e23 - e01;
          }
        x[2] = 0;
      }
  }

  // rows of the adjugate: adj[d][e] = det(J) * (J^-1)[d][e]
  template <int dim>
  __device__ __forceinline__ double
  adjugate(const double (&J)[3][3], double (&adj)[3][3])
  {
    if constexpr (dim == 3)
      {
        // columns c_d = J[.][d]; adj row 0 = c1 x c2, row 1 = c2 x c0,
        // row 2 = c0 x c1
        for (int d = 0; d < 3; ++d)
          {
            const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
            adj[d][0] = J[1][d1] * J[2][d2] - J[2][d1] * J[1][d2];
            adj[d][1] = J[2][d1] * J[0][d2] - J[0][d1] * J[2][d2];
            adj[d][2] = J[0][d1] * J[1][d2] - J[1][d1] * J[0][d2];
          }
        return J[0][0] * adj[0][0] + J[1][0] * adj[0][1] +
               J[2][0] * adj[0][2];
      }
    else
      {
        adj[0][0] = J[1][1];
        adj[0][1] = -J[0][1];
        adj[1][0] = -J[1][0];
        adj[1][1] = J[0][0];
        return J[0][0] * J[1][1] - J[0][1] * J[1][0];
      }
  }


  // sin and cos of pi (a0 + t_i da), i < n
  template <int n>
  __device__ __forceinline__ void
  sincospi_line(double a0, double da, const double *t, double (&s)[n],
                double (&c)[n])
  {
    const double dth = pi * da;
    if (fabs(dth) < 0.25)
      {
        double s0, c0;
        sincospi(a0, &s0, &c0);
#pragma unroll
        for (int i = 0; i < n; ++i)
          {
            const double ph = t[i] * dth, p2 = ph * ph;
            // Taylor series of sin and cos to order 11 and 12
            double sp = -1. / 39916800.;
            sp        = fma(sp, p2, 1. / 362880.);
            sp        = fma(sp, p2, -1. / 5040.);
            sp        = fma(sp, p2, 1. / 120.);
            sp        = fma(sp, p2, -1. / 6.);
            sp        = fma(sp * p2, ph, ph);
            double cp = 1. / 479001600.;
            cp        = fma(cp, p2, -1. / 3628800.);
            cp        = fma(cp, p2, 1. / 40320.);
            cp        = fma(cp, p2, -1. / 720.);
            cp        = fma(cp, p2, 1. / 24.);
            cp        = fma(cp, p2, -0.5);
            cp        = fma(cp, p2, 1.);
            s[i]      = fma(s0, cp, c0 * sp);
            c[i]      = fma(c0, cp, -(s0 * sp));
          }
      }
    else
      {
#pragma unroll
        for (int i = 0; i < n; ++i)
          sincospi(fma(t[i], da, a0), &s[i], &c[i]);
      }
  }

  // velocity field (without time factor) at x0 + t_i dx
  template <int dim, int n>
  __device__ __forceinline__ void
  velocity_line(int kind, const double *vconst, const double (&x0)[3],
                const double (&dx)[3], const double *t, double (&a)[n][3])
  {
    if (kind == vel_cellular_3d)
      {
        double s[3][n], c[3][n];
#pragma unroll
        for (int e = 0; e < 3; ++e)
          sincospi_line<n>(2. * x0[e], 2. * dx[e], t, s[e], c[e]);
#pragma unroll
        for (int i = 0; i < n; ++i)
          {
            a[i][0] = s[1][i] * c[2][i];
            a[i][1] = s[2][i] * c[0][i];
            a[i][2] = s[0][i] * c[1][i];
          }
      }
    else if (kind == vel_vortex || kind == vel_vortex_pert)
      {
        double s[2][n], c[2][n];
#pragma unroll
        for (int e = 0; e < 2; ++e)
          sincospi_line<n>(x0[e], dx[e], t, s[e], c[e]);
#pragma unroll
        for (int i = 0; i < n; ++i)
          {
            const double s2x = 2. * s[0][i] * c[0][i];
            const double s2y = 2. * s[1][i] * c[1][i];
            a[i][0] = s2y * s[0][i] * s[0][i];
            a[i][1] = -(s2x * s[1][i] * s[1][i]);
            a[i][2] = 0.;
          }
        if (kind == vel_vortex_pert)
          {
            double sp[2][n], cp[2][n];
            sincospi_line<n>(80. * (x0[0] + 0.2), 80. * dx[0], t, sp[0],
                             cp[0]);
            sincospi_line<n>(80. * (x0[1] + 0.3), 80. * dx[1], t, sp[1],
                             cp[1]);
#pragma unroll
            for (int i = 0; i < n; ++i)
              {
                a[i][0] += 0.2 * sp[0][i] * cp[1][i];
                a[i][1] -= 0.2 * cp[0][i] * sp[1][i];
              }
          }
#pragma unroll
        for (int i = 0; i < n; ++i)
          {
            a[i][0] *= 2.;
            a[i][1] *= 2.;
          }
      }
    else
      {
#pragma unroll
        for (int i = 0; i < n; ++i)
          for (int e = 0; e < 3; ++e)
            a[i][e] = e < dim ? vconst[e] : 0.;
      }
  }

  // adjugate row d of J: (det J) (J^-1)[d][.]
  template <int dim>
  __device__ __forceinline__ void
  adjugate_row(const double (&J)[3][3], int d, double (&r)[3])
  {
    if constexpr (dim == 3)
      {
        const int d1 = (d + 1) % 3, d2 = (d + 2) % 3;
        r[0] = J[1][d1] * J[2][d2] - J[2][d1] * J[1][d2];
        r[1] = J[2][d1] * J[0][d2] - J[0][d1] * J[2][d2];
        r[2] = J[0][d1] * J[1][d2] - J[1][d1] * J[0][d2];
      }
    else
      {
        if (d == 0)
          {
            r[0] = J[1][1];
            r[1] = -J[0][1];
          }
        else
          {
            r[0] = -J[1][0];
            r[1] = J[0][0];
          }
        r[2] = 0;
      }
  }

  // map of a cell at the two ends (xi_dir = 0, 1) of a coordinate line
  template <int dim, bool cartesian>
  __device__ __forceinline__ void
  line_ends(const OpParams &P, const double *X, const double (&xi)[3],
            int dir, double (&x0)[3], double (&J0)[3][3], double (&x1)[3],
            double (&J1)[3][3])
  {
    double a[3] = {xi[0], xi[1], xi[2]}, b[3] = {xi[0], xi[1], xi[2]};
    a[dir]      = 0.;
    b[dir]      = 1.;
    if constexpr (cartesian)
      {
        for (int e = 0; e < 3; ++e)
          {
            x0[e] = e < dim ? X[e] + P.h[e] * a[e] : 0.;
            x1[e] = e < dim ? X[e] + P.h[e] * b[e] : 0.;
            for (int d = 0; d < 3; ++d)
              J0[e][d] = J1[e][d] = (e == d && e < dim) ? P.h[e] : 0.;
          }
      }
    else
      {
        multilinear_map<dim>(X, a, x0, J0);
        multilinear_map<dim>(X, b, x1, J1);
      }
  }

  // Cell data on the LAST-direction line of thread p: cv[d] = w (adj(J) a)_d
  // and JxW at the n Gauss points, and a.n dA on the two faces the line
  // ends on.
  template <int dim, int n, int LAST, bool cartesian, typename T>
  __device__ __forceinline__ void
  line_cell_data(const OpParams &P, const double *X,
                 const DevBasis<double> *Bd, const int (&q)[3],
                 T (&cv)[3][n], T (&jxw)[n], T (&fan_last)[2])
  {
    double xi[3] = {Bd->xq[q[0]], Bd->xq[q[1]], dim == 3 ? Bd->xq[q[2]] : 0.};
    double wl    = 1.;
    for (int e = 0; e < dim; ++e)
      if (e != LAST)
        wl *= Bd->w[q[e]];
    double x0[3], J0[3][3], x1[3], J1[3][3];
    line_ends<dim, cartesian>(P, X, xi, LAST, x0, J0, x1, J1);
    double dx[3];
    for (int e = 0; e < 3; ++e)
      dx[e] = x1[e] - x0[e];
    double a[n][3];
    velocity_line<dim, n>(P.vel_kind, P.vel_const, x0, dx, Bd->xq, a);
#pragma unroll
    for (int i = 0; i < n; ++i)
      {
        const double t = Bd->xq[i], w = wl * Bd->w[i];
        double       J[3][3], adj[3][3];
        for (int e = 0; e < 3; ++e)
          for (int d = 0; d < 3; ++d)
            J[e][d] = fma(t, J1[e][d] - J0[e][d], J0[e][d]);
        const double det = adjugate<dim>(J, adj);
        for (int d = 0; d < 3; ++d)
          {
            double s = 0;
            for (int e = 0; e < dim; ++e)
              s += adj[d][e] * a[i][e];
            cv[d][i] = d < dim ? (T)(w * s) : (T)0;
          }
        jxw[i] = (T)(w * det);
      }
    // faces at both ends of the line
    double ae[2][3];
    velocity_field<dim>(P.vel_kind, P.vel_const, x0, ae[0]);
    velocity_field<dim>(P.vel_kind, P.vel_const, x1, ae[1]);
    for (int s = 0; s < 2; ++s)
      {
        double r[3];
        adjugate_row<dim>(s ? J1 : J0, LAST, r);
        double an = 0;
        for (int e = 0; e < dim; ++e)
          an += r[e] * ae[s][e];
        fan_last[s] = (T)((s ? 1. : -1.) * wl * an);
      }
  }

  // a.n dA on a face normal to `d` < LAST (side s), along the line in the
  // LAST direction at tangential Gauss point j of the remaining direction;
  // writes the n values to out[i]
  template <int dim, int n, int LAST, bool cartesian, typename T>
  __device__ __forceinline__ void
  face_line_data(const OpParams &P, const double *X,
                 const DevBasis<double> *Bd, int d, int s, int j, T *out)
  {
    double xi[3] = {0, 0, 0};
    double wj    = 1.;
    if (dim == 3)
      {
        const int o = 1 - d; // the remaining direction (d, o < LAST = 2)
        xi[o]       = Bd->xq[j];
        wj          = Bd->w[j];
      }
    xi[d] = s ? 1. : 0.;
    double x0[3], J0[3][3], x1[3], J1[3][3];
    line_ends<dim, cartesian>(P, X, xi, LAST, x0, J0, x1, J1);
    double dx[3];
    for (int e = 0; e < 3; ++e)
      dx[e] = x1[e] - x0[e];
    double a[n][3];
    velocity_line<dim, n>(P.vel_kind, P.vel_const, x0, dx, Bd->xq, a);
#pragma unroll
    for (int i = 0; i < n; ++i)
      {
        const double t = Bd->xq[i];
        double       J[3][3], r[3];
        for (int e = 0; e < 3; ++e)
          for (int dd = 0; dd < 3; ++dd)
            J[e][dd] = fma(t, J1[e][dd] - J0[e][dd], J0[e][dd]);
        adjugate_row<dim>(J, d, r);
        double an = 0;
        for (int e = 0; e < dim; ++e)
          an += r[e] * a[i][e];
        out[i] = (T)((s ? 1. : -1.) * wj * Bd->w[i] * an);
      }
  }
  // JxW = w_q det(J) at the n points of line p of direction dir (the point
  // layout of Lay<dim, n>::idx). On a multilinear cell the Jacobian column
  // along the line is constant and the other columns are linear in the
  // line coordinate, so J is evaluated at the two line ends only.
  template <int dim, int n, int dir, typename T, typename Layout>
  __device__ __forceinline__ void
  line_jxw(const double *X, const double *xq, const double *wq, int p,
           T (&jxw)[n])
  {
    int q[3];
    Layout::template idx<dir>(p, 0, q);
    double xi[3] = {xq[q[0]], xq[q[1]], dim == 3 ? xq[q[2]] : 0.};
    double wl    = 1.;
#pragma unroll
    for (int e = 0; e < dim; ++e)
      if (e != dir)
        wl *= wq[q[e]];
    double x[3], J0[3][3], J1[3][3];
    xi[dir] = 0.;
    multilinear_map<dim>(X, xi, x, J0);
    xi[dir] = 1.;
    multilinear_map<dim>(X, xi, x, J1);
#pragma unroll
    for (int i = 0; i < n; ++i)
      {
        const double t = xq[i];
        double       J[3][3];
#pragma unroll
        for (int a = 0; a < dim; ++a)
#pragma unroll
          for (int b = 0; b < dim; ++b)
            J[a][b] = fma(t, J1[a][b] - J0[a][b], J0[a][b]);
        double det;
        if constexpr (dim == 3)
          det = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) +
                J[1][0] * (J[2][1] * J[0][2] - J[0][1] * J[2][2]) +
                J[2][0] * (J[0][1] * J[1][2] - J[1][1] * J[0][2]);
        else
          det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        jxw[i] = (T)(wl * wq[i] * det);
      }
  }

  // 1/JxW at the Gauss points of line p of direction dir, as line_jxw.
  // The reciprocal of det J is the hardware estimate refined by two
  // Newton steps (relative error ~1 ulp; det J > 0 and well scaled on a
  // valid mesh).
  template <int dim, int n, int dir, typename T, typename Layout>
  __device__ __forceinline__ void
  line_inv_jxw(const double *X, const double *xq, const double *wq,
               const double *winv, int p, T (&ij)[n])
  {
    int q[3];
    Layout::template idx<dir>(p, 0, q);
    double xi[3] = {xq[q[0]], xq[q[1]], dim == 3 ? xq[q[2]] : 0.};
    double wl    = 1.;
#pragma unroll
    for (int e = 0; e < dim; ++e)
      if (e != dir)
        wl *= wq[q[e]];
    const double rwl = 1. / wl;
    double       x[3], J0[3][3], J1[3][3];
    xi[dir] = 0.;
    multilinear_map<dim>(X, xi, x, J0);
    xi[dir] = 1.;
    multilinear_map<dim>(X, xi, x, J1);
#pragma unroll
    for (int i = 0; i < n; ++i)
      {
        const double t = xq[i];
        double       J[3][3];
#pragma unroll
        for (int a = 0; a < dim; ++a)
#pragma unroll
          for (int b = 0; b < dim; ++b)
            J[a][b] = fma(t, J1[a][b] - J0[a][b], J0[a][b]);
        double det;
        if constexpr (dim == 3)
          det = J[0][0] * (J[1][1] * J[2][2] - J[2][1] * J[1][2]) +
                J[1][0] * (J[2][1] * J[0][2] - J[0][1] * J[2][2]) +
                J[2][0] * (J[0][1] * J[1][2] - J[1][1] * J[0][2]);
        else
          det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        double r = __builtin_amdgcn_rcp(det);
        r        = fma(r, fma(-det, r, 1.), r);
        r        = fma(r, fma(-det, r, 1.), r);
        ij[i]    = (T)(rwl * winv[i] * r);
      }
  }
} // namespace dgops
