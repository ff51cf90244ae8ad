// Basis changes and the per-cell update of the batched cell solver of the
// IRK block-Jacobi preconditioner (advection_solver_irk_real.cc:2108-2126:
// CellwiseOperator::transform_to_collocation, one right-preconditioned
// GMRES step per cell, transform_from_collocation), in the line-per-thread
// layout of the operator kernel.
#pragma once

#include "op_kernel.hpp"

namespace dgops
{
  // dst = (S^-T x S^-T x S^-T) src on every cell and stage block
  // (blockIdx.y): a weak-form vector tested with the Gauss-Lobatto nodal
  // functions -> the same vector tested with the Lagrange functions of the
  // Gauss points (phi_i = sum_q S[q][i] ell_q), stored lexicographically.
  // With the cell operator in that basis, A_c = S^-T A S^-1, the cell
  // inverse is A^-1 = S^-1 A_c^-1 S^-T.
  template <int dim, int n, typename T>
  __global__ void __launch_bounds__(128)
  to_collocation_kernel(int64_t n_cells, int64_t block,
                        const DevBasis<T> *__restrict__ Bp,
                        const T *__restrict__ src, T *__restrict__ dst)
  {
    using L            = Lay<dim, n>;
    constexpr int nl   = L::nl;
    constexpr int nd   = L::nd;
    constexpr int CPB  = cells_per_block<dim, n>();
    constexpr int STR  = L::stride;
    constexpr int LAST = dim - 1;
    __shared__ T  s_a[CPB * STR];

    const DevBasis<T> &B      = *Bp;
    const int          tid    = threadIdx.x;
    const int          lc     = tid / nl;
    const int          p      = tid - lc * nl;
    const int64_t      cell   = (int64_t)xcd_block(blockIdx.x, gridDim.x) * CPB + lc;
    const bool         active = (lc < CPB) && (cell < n_cells);
    T                 *a_l    = s_a + (lc < CPB ? lc : 0) * STR;
    const T           *in     = src + blockIdx.y * block + cell * nd;
    T                 *out    = dst + blockIdx.y * block + cell * nd;

    T r[n], t[n];
#pragma unroll
    for (int i = 0; i < n; ++i)
      r[i] = active ? in[L::template off<0>(p, i)] : (T)0;
    eo_apply<n, false, true, false>(B.eo_Sinv, r, t);
    lds_write_line<dim, n, 0>(a_l, p, t);
    __syncthreads();
    lds_read_line<dim, n, 1>(a_l, p, t);
    eo_apply<n, false, true, false>(B.eo_Sinv, t, r);
    if constexpr (dim == 3)
      {
        __syncthreads();
        lds_write_line<dim, n, 1>(a_l, p, r);
        __syncthreads();
        lds_read_line<dim, n, 2>(a_l, p, t);
        eo_apply<n, false, true, false>(B.eo_Sinv, t, r);
      }
    if (active)
#pragma unroll
      for (int i = 0; i < n; ++i)
        out[L::template off<LAST>(p, i)] = r[i];
  }

  // One GMRES step per cell with right preconditioning from x0 = 0, for
  // the NS stage blocks of the cell together (deal.II's batched mode: one
  // solve per cell): with r the right-hand side, y = P^-1 r and z = A y,
  // the minimizer of |r - alpha z| over the cell is
  //   alpha = <z, r> / <z, z>,   x = alpha y.
  // Then dst = (S^-1 x S^-1 x S^-1) x, back to Gauss-Lobatto nodal values.
  // r, y, z hold values at the Gauss points, lexicographic.
  template <int dim, int n, int NS, typename T>
  __global__ void __launch_bounds__(128)
  cell_gmres_update_kernel(int64_t n_cells, int64_t block,
                           const DevBasis<T> *__restrict__ Bp,
                           const T *__restrict__ r, const T *__restrict__ y,
                           const T *__restrict__ z, T *__restrict__ dst)
  {
    using L            = Lay<dim, n>;
    constexpr int nl   = L::nl;
    constexpr int nd   = L::nd;
    constexpr int CPB  = cells_per_block<dim, n>();
    constexpr int STR  = L::stride;
    constexpr int LAST = dim - 1;
    __shared__ T      s_a[CPB * STR];
    __shared__ double s_zr[CPB * nl], s_zz[CPB * nl];
    __shared__ double s_alpha[CPB];

    const DevBasis<T> &B      = *Bp;
    const int          tid    = threadIdx.x;
    const int          lc     = tid / nl;
    const int          p      = tid - lc * nl;
    const int64_t      cell   = (int64_t)xcd_block(blockIdx.x, gridDim.x) * CPB + lc;
    const bool         active = (lc < CPB) && (cell < n_cells);
    const int          lcs    = lc < CPB ? lc : 0;
    T                 *a_l    = s_a + lcs * STR;

    // the cell's two inner products, reduced over its lines in LDS
    double zr = 0, zz = 0;
    if (active)
      for (int c = 0; c < NS; ++c)
#pragma unroll
        for (int i = 0; i < n; ++i)
          {
            const int64_t k  = c * block + cell * nd + L::template off<LAST>(p, i);
            const double  zk = (double)z[k];
            zr += zk * (double)r[k];
            zz += zk * zk;
          }
    if (lc < CPB)
      {
        s_zr[lc * nl + p] = zr;
        s_zz[lc * nl + p] = zz;
      }
    __syncthreads();
    if (lc < CPB && p == 0)
      {
        double a = 0, b = 0;
        for (int j = 0; j < nl; ++j)
          {
            a += s_zr[lc * nl + j];
            b += s_zz[lc * nl + j];
          }
        s_alpha[lc] = b > 0 ? a / b : 0.;
      }
    __syncthreads();
    const T alpha = (T)s_alpha[lcs];

    for (int c = 0; c < NS; ++c)
      {
        T u[n], t[n];
#pragma unroll
        for (int i = 0; i < n; ++i)
          u[i] = active ? alpha * y[c * block + cell * nd +
                                    L::template off<LAST>(p, i)] :
                          (T)0;
        eo_apply<n, false, false, false>(B.eo_Sinv, u, t);
        if constexpr (dim == 3)
          {
            lds_write_line<dim, n, 2>(a_l, p, t);
            __syncthreads();
            lds_read_line<dim, n, 1>(a_l, p, u);
            eo_apply<n, false, false, false>(B.eo_Sinv, u, t);
            __syncthreads();
          }
        lds_write_line<dim, n, 1>(a_l, p, t);
        __syncthreads();
        lds_read_line<dim, n, 0>(a_l, p, u);
        eo_apply<n, false, false, false>(B.eo_Sinv, u, t);
        if (active)
#pragma unroll
          for (int i = 0; i < n; ++i)
            dst[c * block + cell * nd + L::template off<0>(p, i)] = t[i];
        __syncthreads();
      }
  }
} // namespace dgops
