// Cellwise fast-diagonalization (FDM) preconditioner for gfx950: the exact
// inverse of the separable part of the cell operator
//   detJ [ (Lambda_stage / dt) (x) M(x)M(x)M + I (x) sum_d a_d M..K_d..M ]
// (block Jacobi: the inflow neighbour coupling is dropped), as in
// CellwisePreconditionerFDM (advection_solver_irk_real.cc:807-1199; single
// stage advection_solver_implicit_real.cc:539-813):
//   1. stage transform T^-1 (IRK)                          :950-985
//   2. d sweeps with V^-1 M^-1                             :997-1009
//   3. pointwise inverse on the real-pair blocks           :1019-1052
//   4. d sweeps with V                                     :1055-1065
//   5. stage transform T                                   :1068-1110
// Per cell the 1D eigenbasis of each direction is the one of the sign of
// the cell's mean reference-space velocity (:838-844).
//
// Step 3 diagonalizes every 2x2 real-pair block [[a, b], [-b, a]] with
// U = [[1, 1], [i, -i]] (eigenvalues lambda, conj(lambda)) and divides by
// sigma + sum_d a_d lambda_d in complex arithmetic; the same holds for the
// stage pairs of A^-1. Lines are owned by threads as in op_kernel.hpp.
#pragma once

#include "op_kernel.hpp"

namespace dgops
{
  template <typename T>
  struct FdmDev
  {
    T   VinvMinv[2][max_n][max_n]; // [sign][i][j]
    T   V[2][max_n][max_n];
    T   lre[2][max_n], lim[2][max_n];
    int kind[max_n];               // 0 real, 1 first of pair, 2 second
    int n_blocks;                  // blocks per direction
    int block_col[max_n];          // first column of block b
    int block_size[max_n];         // 1 or 2
    // stages (Radau IIA A^-1 = T Lambda T^-1), s <= 5
    T   Tinv[5][5], Tm[5][5];
    T   mre[5], mim[5];
    int n_stage_blocks;
    int stage_block_col[5];
    int stage_block_size[5];
  };

  struct FdmParams
  {
    int64_t       n_cells;
    const double *cell_velocity; // [cell][dim], reference space, t = 0
    const double *cell_invdet;   // [cell]
    double        vel_scale;     // time factor
    double        inv_dt;
    int64_t       block;         // stage block size (dofs)
  };

  template <int dim, int n, typename T>
  __device__ __forceinline__ void
  fdm_block(T *data, const FdmDev<T> &F, const int (&sg)[3],
            const T (&abar)[3], T invdet, T inv_dt, int sb, int bx, int by,
            int bz)
  {
    constexpr int STR = Lay<dim, n>::stride;
    // factors: 0 stage, 1 x, 2 y, 3 z
    int fcol[4], fsize[4];
    fcol[0]  = F.stage_block_col[sb];
    fsize[0] = F.stage_block_size[sb];
    const int bb[3] = {bx, by, bz};
    for (int d = 0; d < 3; ++d)
      {
        if (d < dim)
          {
            fcol[1 + d]  = F.block_col[bb[d]];
            fsize[1 + d] = F.block_size[bb[d]];
          }
        else
          {
            fcol[1 + d]  = 0;
            fsize[1 + d] = 1;
          }
      }
    int bitf[4], nbits = 0;
    for (int f = 0; f < 4; ++f)
      if (fsize[f] == 2)
        bitf[nbits++] = f;
    const int ne = 1 << nbits;
    T         re[16], im[16];
    for (int e = 0; e < ne; ++e)
      {
        int off[4] = {0, 0, 0, 0};
        for (int b = 0; b < nbits; ++b)
          off[bitf[b]] = (e >> b) & 1;
        const int js = fcol[0] + off[0];
        const int ix = fcol[1] + off[1], iy = fcol[2] + off[2],
                  iz = fcol[3] + off[3];
        re[e] = data[js * STR + ix + n * (iy + n * iz)];
        im[e] = 0;
      }
    // U^-1 on every pair bit: (x0, x1) -> (1/2 (x0 - i x1), 1/2 (x0 + i x1))
    for (int b = 0; b < nbits; ++b)
      for (int e = 0; e < ne; ++e)
        if (!((e >> b) & 1))
          {
            const int e1 = e | (1 << b);
            const T   r0 = re[e], i0 = im[e], r1 = re[e1], i1 = im[e1];
            re[e]        = (T)0.5 * (r0 + i1);
            im[e]        = (T)0.5 * (i0 - r1);
            re[e1]       = (T)0.5 * (r0 - i1);
            im[e1]       = (T)0.5 * (i0 + r1);
          }
    // divide by sigma + sum_d a_d lambda_d (bit 1: conjugate eigenvalue)
    for (int e = 0; e < ne; ++e)
      {
        int off[4] = {0, 0, 0, 0};
        for (int b = 0; b < nbits; ++b)
          off[bitf[b]] = (e >> b) & 1;
        const int js = fcol[0];
        T         lr = F.mre[js] * inv_dt;
        T         li = (fsize[0] == 2 ? (off[0] ? -F.mim[js] : F.mim[js]) :
                                        (T)0) *
               inv_dt;
        for (int d = 0; d < dim; ++d)
          {
            const int c = fcol[1 + d];
            lr += abar[d] * F.lre[sg[d]][c];
            if (fsize[1 + d] == 2)
              li += abar[d] * (off[1 + d] ? -F.lim[sg[d]][c] : F.lim[sg[d]][c]);
          }
        const T den = lr * lr + li * li;
        const T sr = invdet * lr / den, si = -invdet * li / den;
        const T r = re[e], i = im[e];
        re[e]     = r * sr - i * si;
        im[e]     = r * si + i * sr;
      }
    // U on every pair bit: (y+, y-) -> (y+ + y-, i (y+ - y-))
    for (int b = 0; b < nbits; ++b)
      for (int e = 0; e < ne; ++e)
        if (!((e >> b) & 1))
          {
            const int e1 = e | (1 << b);
            const T   r0 = re[e], i0 = im[e], r1 = re[e1], i1 = im[e1];
            re[e]        = r0 + r1;
            im[e]        = i0 + i1;
            re[e1]       = -(i0 - i1);
            im[e1]       = r0 - r1;
          }
    for (int e = 0; e < ne; ++e)
      {
        int off[4] = {0, 0, 0, 0};
        for (int b = 0; b < nbits; ++b)
          off[bitf[b]] = (e >> b) & 1;
        const int js = fcol[0] + off[0];
        const int ix = fcol[1] + off[1], iy = fcol[2] + off[2],
                  iz = fcol[3] + off[3];
        data[js * STR + ix + n * (iy + n * iz)] = re[e];
      }
  }

  template <int dim, int n, int NS, typename Tc, typename Tin, typename Tout>
  __global__ void __launch_bounds__(256)
  fdm_kernel(const FdmParams P, const FdmDev<Tc> *__restrict__ Fp,
             const Tin *__restrict__ src, Tout *__restrict__ dst)
  {
    using L           = Lay<dim, n>;
    constexpr int nl  = L::nl;
    constexpr int nd  = L::nd;
    constexpr int CPB = cells_per_block<dim, n>();
    constexpr int STR = L::stride;
    constexpr int LAST = dim - 1;

    __shared__ Tc s_data[CPB * NS * STR];
    __shared__ Tc s_t[CPB * STR];

    const FdmDev<Tc> &F    = *Fp;
    const int         tid  = threadIdx.x;
    const int         lc   = tid / nl;
    const int         p    = tid - lc * nl;
    const int64_t     cell = (int64_t)blockIdx.x * CPB + lc;
    const bool        active = (lc < CPB) && (cell < P.n_cells);
    const int         lcs    = lc < CPB ? lc : 0;
    Tc               *data   = s_data + lcs * NS * STR;
    Tc               *tb     = s_t + lcs * STR;

    int sg[3] = {0, 0, 0};
    Tc  abar[3] = {0, 0, 0};
    Tc  invdet  = 1;
    if (active)
      {
        for (int d = 0; d < dim; ++d)
          {
            const double a = P.cell_velocity[cell * dim + d] * P.vel_scale;
            abar[d]        = (Tc)a;
            sg[d]          = a < 0. ? 1 : 0;
          }
        invdet = (Tc)P.cell_invdet[cell];
      }

    // 1-2: load, stage transform, V^-1 M^-1 sweeps x, y, (z)
    Tc w[NS][n];
    {
      Tc r[NS][n];
#pragma unroll
      for (int t = 0; t < NS; ++t)
#pragma unroll
        for (int i = 0; i < n; ++i)
          r[t][i] = active ? (Tc)src[t * P.block + cell * nd +
                                     L::template off<0>(p, i)] :
                             (Tc)0;
#pragma unroll
      for (int j = 0; j < NS; ++j)
#pragma unroll
        for (int i = 0; i < n; ++i)
          {
            Tc s = 0;
#pragma unroll
            for (int t = 0; t < NS; ++t)
              s = fma(F.Tinv[j][t], r[t][i], s);
            w[j][i] = NS == 1 ? r[0][i] : s;
          }
    }
#pragma unroll
    for (int j = 0; j < NS; ++j)
      {
        Tc t[n];
        mat_apply<n>(F.VinvMinv[sg[0]], w[j], t);
        lds_write_line<dim, n, 0>(tb, p, t);
        __syncthreads();
        lds_read_line<dim, n, 1>(tb, p, t);
        mat_apply<n>(F.VinvMinv[sg[1]], t, w[j]);
        if constexpr (dim == 3)
          {
            __syncthreads();
            lds_write_line<dim, n, 1>(tb, p, w[j]);
            __syncthreads();
            lds_read_line<dim, n, 2>(tb, p, t);
            mat_apply<n>(F.VinvMinv[sg[2]], t, w[j]);
          }
        lds_write_line<dim, n, LAST>(data + j * STR, p, w[j]);
        __syncthreads();
      }

    // 3: pointwise inverse on the blocks of this cell
    {
      const int nbd  = F.n_blocks;
      const int nspb = dim == 3 ? nbd * nbd * nbd : nbd * nbd;
      const int ntask = F.n_stage_blocks * nspb;
      if (active)
        for (int task = p; task < ntask; task += nl)
          {
            const int sb = task / nspb;
            int       rb = task - sb * nspb;
            const int bx = rb % nbd;
            rb /= nbd;
            const int by = rb % nbd;
            const int bz = dim == 3 ? rb / nbd : 0;
            fdm_block<dim, n, Tc>(data, F, sg, abar, invdet, (Tc)P.inv_dt, sb,
                                  bx, by, bz);
          }
      __syncthreads();
    }

    // 4: V sweeps LAST, ..., x
#pragma unroll
    for (int j = 0; j < NS; ++j)
      {
        Tc t[n];
        lds_read_line<dim, n, LAST>(data + j * STR, p, t);
        mat_apply<n>(F.V[sg[LAST]], t, w[j]);
        if constexpr (dim == 3)
          {
            lds_write_line<dim, n, 2>(tb, p, w[j]);
            __syncthreads();
            lds_read_line<dim, n, 1>(tb, p, t);
            mat_apply<n>(F.V[sg[1]], t, w[j]);
            __syncthreads();
            lds_write_line<dim, n, 1>(tb, p, w[j]);
          }
        else
          lds_write_line<dim, n, 1>(tb, p, w[j]);
        __syncthreads();
        lds_read_line<dim, n, 0>(tb, p, t);
        mat_apply<n>(F.V[sg[0]], t, w[j]);
        __syncthreads();
      }

    // 5: back stage transform and store (x-lines)
    if (active)
#pragma unroll
      for (int t = 0; t < NS; ++t)
#pragma unroll
        for (int i = 0; i < n; ++i)
          {
            Tc s = 0;
#pragma unroll
            for (int j = 0; j < NS; ++j)
              s = fma(F.Tm[t][j], w[j][i], s);
            dst[t * P.block + cell * nd + L::template off<0>(p, i)] =
              (Tout)(NS == 1 ? w[0][i] : s);
          }
  }
} // namespace dgops
