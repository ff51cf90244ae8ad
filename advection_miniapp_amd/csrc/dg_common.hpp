// Plain-data structures shared by the host code and the HIP kernels.
#pragma once

#include <cstdint>

namespace dgops
{
  // analytic velocity fields (velocity = factor * field(x))
  enum VelocityKind : int
  {
    vel_constant     = 0, // vel_const
    vel_cellular_3d  = 1, // (sin2πy cos2πz, sin2πz cos2πx, sin2πx cos2πy)
    vel_vortex       = 2, // advection_solver_variable.cc:163-176
    vel_vortex_pert  = 3, // advection_solver_implicit_real.cc:168-184
  };

  // analytic solutions / initial data g(x,t)
  enum SolutionKind : int
  {
    sol_zero         = 0,
    sol_wave         = 1, // advection_solver.cc:145-158 (wave number 4)
    sol_sine_product = 2, // prod_d sin(w (x_d - a_d t)) * exp(-nu d w^2 t)
    sol_gaussian     = 3, // exp(-|x-c|^2 / width^2)
    sol_polynomial   = 4, // (1 + x + 2y + 3z - (a.(1,2,3)) t)^p
    sol_two_pulses   = 5, // advection_solver_variable.cc:141-147
    sol_burgers_sine = 6, // burgers_solver.cc:109-117, component 0 only
  };

  // parameters of one operator evaluation, passed by value to kernels
  struct OpParams
  {
    int64_t       n_cells;
    const int    *neighbors;     // [cell][2 dim]
    const double *cell_vertices; // [cell][2^dim][dim]
    double        h[3];          // cell size when `cartesian`
    int           cartesian;
    int           vel_kind;
    double        vel_const[3];
    double        vel_factor[8]; // per component/stage time factor
    double        alpha;         // flux_alpha
    double        skew;          // factor_skew
    double        sign;          // -1: explicit right-hand side, +1: operator
    double        mass[8][8];    // mass coupling factors [c][b] (implicit)
    int           has_mass;
    int           sol_kind;      // Dirichlet data
    double        sol_time;
    double        sol_param[8];
    double        viscosity;
    double        penalty;
    // >= 0: the mesh is one periodic 3D coarse cell refined morton_bits
    // times, so the cell index is the Morton code of (i, j, k) (x in bit
    // 3b) and the face neighbours follow from it; -1: read `neighbors`
    int           morton_bits;
    // geometry tables of the tabulated-geometry kernels (scalar type of
    // the kernel): per cell [dim + 1][nd] = w adj(J) a (dim rows) and
    // 1/JxW at the Gauss points, each row in the line order of the
    // direction it is read along (row d along d, row dim along x: point i
    // of line p at i * nl + p; JxW of the implicit mass term is recomputed
    // from the corners along the lines); per
    // cell [dim][nl] = w (adj(J) a)_d on the face at xi_d = 0 (normal
    // along +xi_d); per boundary face [nl][1 + dim] (double) = outward
    // a . n dA and the physical point
    const void   *tab_cell;
    const void   *tab_face;
    const double *tab_bnd;
    const int    *bnd_index; // [cell][2 dim] boundary face id, -1 inside
    const double *bnd_value; // [boundary face][nl] Dirichlet data at sol_time
  };

  // LSRK stage update (advection_solver.cc:728-751)
  struct StageParams
  {
    double b_dt;
    double a_dt;
    int    last; // a_dt == 0: only the solution update
  };
} // namespace dgops
