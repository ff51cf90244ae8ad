// Analytic velocity fields and solutions, usable on host and device.
#pragma once

#include <cmath>

#include "dg_common.hpp"

#ifdef __HIPCC__
#  define DG_HD __host__ __device__
#else
#  define DG_HD
#endif

namespace dgops
{
  constexpr double pi = 3.14159265358979323846;

  DG_HD inline void
  sincospi_(double x, double *s, double *c)
  {
#ifdef __HIP_DEVICE_COMPILE__
    sincospi(x, s, c);
#else
    *s = std::sin(pi * x);
    *c = std::cos(pi * x);
#endif
  }

  // velocity field without its time factor
  template <int dim>
  DG_HD inline void
  velocity_field(int kind, const double *vconst, const double *x, double *a)
  {
    if (kind == vel_cellular_3d)
      {
        double sx, cx, sy, cy, sz = 0., cz = 1.;
        sincospi_(2. * x[0], &sx, &cx);
        sincospi_(2. * x[1], &sy, &cy);
        if (dim == 3)
          sincospi_(2. * x[2], &sz, &cz);
        a[0] = sy * cz;
        a[1] = sz * cx;
        if (dim == 3)
          a[2] = sx * cy;
      }
    else if (kind == vel_vortex || kind == vel_vortex_pert)
      {
        double s2x, c2x, s2y, c2y, s1x, c1x, s1y, c1y;
        sincospi_(2. * x[0], &s2x, &c2x);
        sincospi_(2. * x[1], &s2y, &c2y);
        sincospi_(x[0], &s1x, &c1x);
        sincospi_(x[1], &s1y, &c1y);
        double a0 = s2y * s1x * s1x;
        double a1 = -(s2x * s1y * s1y);
        if (kind == vel_vortex_pert)
          {
            double spx, cpx, spy, cpy;
            sincospi_(80. * (x[0] + 0.2), &spx, &cpx);
            sincospi_(80. * (x[1] + 0.3), &spy, &cpy);
            a0 += 0.2 * spx * cpy;
            a1 -= 0.2 * cpx * spy;
          }
        a[0] = 2. * a0;
        a[1] = 2. * a1;
        if (dim == 3)
          a[2] = 0.;
      }
    else
      {
        for (int d = 0; d < dim; ++d)
          a[d] = vconst[d];
      }
  }

  template <int dim>
  DG_HD inline double
  analytic_solution(int           kind,
                    const double *p,
                    double        t,
                    const double *a,
                    const double *x,
                    int           component = 0)
  {
    switch (kind)
      {
        case sol_wave:
          {
            double r = std::sin(4. * pi * (x[0] - t * a[0]));
            for (int d = 1; d < dim; ++d)
              r *= std::cos(4. * pi * (x[d] - t * a[d]));
            return r;
          }
        case sol_sine_product:
          {
            double r = 1.;
            for (int d = 0; d < dim; ++d)
              r *= std::sin(p[0] * (x[d] - t * a[d]));
            if (p[1] != 0.)
              r *= std::exp(-p[1] * dim * p[0] * p[0] * t);
            return r;
          }
        case sol_gaussian:
          {
            double r2 = 0;
            for (int d = 0; d < dim; ++d)
              r2 += (x[d] - p[d]) * (x[d] - p[d]);
            return std::exp(-r2 / (p[3] * p[3]));
          }
        case sol_polynomial:
          {
            double s = 1.;
            for (int d = 0; d < dim; ++d)
              s += (d + 1) * (x[d] - t * a[d]);
            double r = 1.;
            for (int e = 0; e < (int)p[0]; ++e)
              r *= s;
            return r;
          }
        case sol_two_pulses:
          return std::exp(-400. * ((x[0] - 0.5) * (x[0] - 0.5) +
                                   (x[1] - 0.75) * (x[1] - 0.75))) +
                 std::exp(-100. * ((x[0] - 0.5) * (x[0] - 0.5) +
                                   (x[1] - 0.25) * (x[1] - 0.25)));
        case sol_burgers_sine:
          return component == 0 ? -std::sin(pi * x[0]) : 0.;
        default:
          return 0.;
      }
  }
} // namespace dgops
