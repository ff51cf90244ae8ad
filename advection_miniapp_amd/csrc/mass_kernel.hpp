// Cellwise inverse mass matrix, M^-1 = S^-1 diag(1/JxW) S^-T per cell
// (CellwiseInverseMassMatrix::apply, advection_solver.cc:635-657), in the
// line-per-thread layout of the operator kernel: three S^-T sweeps, the
// pointwise 1/JxW on LAST-direction lines (recomputed from the cell's
// corners), three S^-1 sweeps. In place; every workgroup reads and writes
// only its own cells.
#pragma once

#include "op_kernel.hpp"

namespace dgops
{
  template <int dim, int n, typename T>
  __global__ void __launch_bounds__(128)
  inverse_mass_line_kernel(const OpParams                       P,
                           const DevBasis<T> *__restrict__      Bp,
                           const DevBasis<double> *__restrict__ Bd,
                           T *__restrict__                      vec)
  {
    using L            = Lay<dim, n>;
    constexpr int nl   = L::nl;
    constexpr int nd   = L::nd;
    constexpr int CPB  = cells_per_block<dim, n>();
    constexpr int STR  = L::stride;
    constexpr int NV   = 1 << dim;
    constexpr int LAST = dim - 1;
    __shared__ T      s_a[CPB * STR];
    __shared__ double s_X[CPB * NV * dim];

    const DevBasis<T> &B      = *Bp;
    const unsigned     block  = xcd_block(blockIdx.x, gridDim.x);
    const int          tid    = threadIdx.x;
    const int          lc     = tid / nl;
    const int          p      = tid - lc * nl;
    const int64_t      cell   = (int64_t)block * CPB + lc;
    const bool         active = (lc < CPB) && (cell < P.n_cells);
    const int          lcs    = lc < CPB ? lc : 0;
    T                 *a_l    = s_a + lcs * STR;
    double            *X      = s_X + lcs * NV * dim;

    T r[n], t[n];
#pragma unroll
    for (int i = 0; i < n; ++i)
      r[i] = active ? vec[cell * nd + L::template off<0>(p, i)] : (T)0;
    if (active)
      for (int e = p; e < NV * dim; e += nl)
        X[e] = P.cell_vertices[cell * NV * dim + e];

    // S^-T sweeps: x (registers), then y (and z) through LDS
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

    // 1/JxW at the Gauss points of this thread's LAST-direction line
    T ij[n];
    if (active)
      line_inv_jxw<dim, n, LAST, T, L>(X, Bd->xq, Bd->w, Bd->winv, p, ij);
    else
#pragma unroll
      for (int i = 0; i < n; ++i)
        ij[i] = 0;
#pragma unroll
    for (int i = 0; i < n; ++i)
      r[i] *= ij[i];

    // S^-1 sweeps back: LAST (registers), then through LDS to x
    eo_apply<n, false, false, false>(B.eo_Sinv, r, t);
    if constexpr (dim == 3)
      {
        __syncthreads();
        lds_write_line<dim, n, 2>(a_l, p, t);
        __syncthreads();
        lds_read_line<dim, n, 1>(a_l, p, r);
        eo_apply<n, false, false, false>(B.eo_Sinv, r, t);
      }
    __syncthreads();
    lds_write_line<dim, n, 1>(a_l, p, t);
    __syncthreads();
    lds_read_line<dim, n, 0>(a_l, p, r);
    eo_apply<n, false, false, false>(B.eo_Sinv, r, t);
    if (active)
#pragma unroll
      for (int i = 0; i < n; ++i)
        vec[cell * nd + L::template off<0>(p, i)] = t[i];
  }
} // namespace dgops
