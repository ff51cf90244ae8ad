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
// The reference keeps one 1D eigenbasis per velocity sign and picks the one
// of the sign of the cell's mean reference-space velocity (:838-844). The
// upwind matrices of the two signs are mirror images,
//   M^-1 K(-) = -R M^-1 K(+) R,   R M R = M,   R = line reversal,
// so a direction with a < 0 uses the eigenbasis of the + sign on the
// reversed line and the eigenvalues of a M^-1 K(-) are |a| lambda(+). The
// reversal is a permutation of the line's points: it is folded into the
// per-lane addresses of the loads, LDS accesses and stores, and the 1D
// tables stay uniform across the wavefront.
//
// Step 3 diagonalizes every 2x2 real-pair block [[a, b], [-b, a]] with
// U = [[1, 1], [i, -i]] (eigenvalues lambda, conj(lambda)) and divides by
// sigma + sum_d |a_d| lambda_d in complex arithmetic; the same holds for
// the stage pairs of A^-1. Lines are owned by threads as in op_kernel.hpp.
#pragma once

#include "op_kernel.hpp"

namespace dgops
{
  // step-3 tasks per cell: at most 2 stage blocks (s <= 4) times 4^3
  // direction blocks (n <= 8)
  constexpr int max_fdm_tasks = 128;

  // one step-3 task: a block of real-pair-coupled points of one cell
  template <typename T>
  struct FdmTask
  {
    // bits 0-1: stage block column, bit 2: stage pair; direction d: bits
    // 3+4d .. 5+4d block column, bit 6+4d pair
    int code;
    T   re[4]; // eigenvalue real parts: stage (times 1/dt), directions
    T   im[4]; // (times |a_d|); imaginary parts, 0 for real blocks
  };

  template <typename T>
  struct FdmDev
  {
    // 1D eigenbasis of M^-1 K(+) (velocity sign +)
    T   VinvMinv[max_n][max_n]; // [i][j]
    T   V[max_n][max_n];
    T   lre[max_n], lim[max_n];
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
    // step-3 tasks of one cell sorted by their number of paired factors
    // NB: tasks[task_start[NB] .. task_start[NB + 1])
    int        task_start[6];
    FdmTask<T> tasks[max_fdm_tasks];
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

  // point i of a line, read or written mirrored when `rev`
  template <int n>
  __device__ __forceinline__ int
  mirror(int i, bool rev)
  {
    return rev ? n - 1 - i : i;
  }

  template <int dim, int n, int dir, typename T>
  __device__ __forceinline__ void
  lds_read_line_m(const T *buf, int p, bool rev, T (&v)[n])
  {
#pragma unroll
    for (int i = 0; i < n; ++i)
      v[i] = buf[Lay<dim, n>::template off<dir>(p, mirror<n>(i, rev))];
  }

  template <int dim, int n, int dir, typename T>
  __device__ __forceinline__ void
  lds_write_line_m(T *buf, int p, bool rev, const T (&v)[n])
  {
#pragma unroll
    for (int i = 0; i < n; ++i)
      buf[Lay<dim, n>::template off<dir>(p, mirror<n>(i, rev))] = v[i];
  }

  // inverse of one block of 2^NB real-pair-coupled points (NB paired
  // factors among stage, x, y, z): load, U^-1 on every paired factor,
  // complex division by the eigenvalue sum, U, store. `delta[b]` is the
  // storage distance to the partner along paired factor b, lr the real
  // part of the eigenvalue sum and li[b] the imaginary part of paired
  // factor b (conjugated when the point's bit b is set).
  template <int NB, typename T>
  __device__ __forceinline__ void
  fdm_block(T *data, int base, const int (&delta)[4], T lr,
            const T (&li)[4], T invdet)
  {
    constexpr int NE = 1 << NB;
    T             re[NE], im[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e)
      {
        int a = base;
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if ((e >> b) & 1)
            a += delta[b];
        re[e] = data[a];
        im[e] = 0;
      }
    // U^-1 on every pair bit: (x0, x1) -> (1/2 (x0 - i x1), 1/2 (x0 + i x1))
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int e = 0; e < NE; ++e)
        if (!((e >> b) & 1))
          {
            const int e1 = e | (1 << b);
            const T   r0 = re[e], i0 = im[e], r1 = re[e1], i1 = im[e1];
            re[e]        = (T)0.5 * (r0 + i1);
            im[e]        = (T)0.5 * (i0 - r1);
            re[e1]       = (T)0.5 * (r0 - i1);
            im[e1]       = (T)0.5 * (i0 + r1);
          }
    // times invdet / (lr + i lim_e), lim_e = sum_b +-li[b]: the complement
    // point (every pair bit flipped) has -lim_e and shares the denominator,
    // so one division serves both
    constexpr int NH = NB == 0 ? 1 : NE / 2;
#pragma unroll
    for (int e = 0; e < NH; ++e)
      {
        T lim = 0;
#pragma unroll
        for (int b = 0; b < NB; ++b)
          lim += ((e >> b) & 1) ? -li[b] : li[b];
        const T q  = invdet / (lr * lr + lim * lim);
        const T sr = lr * q, si = -lim * q;
        {
          const T r = re[e], i = im[e];
          re[e]     = r * sr - i * si;
          im[e]     = r * si + i * sr;
        }
        if constexpr (NB > 0)
          {
            const int ec = e ^ (NE - 1);
            const T   r  = re[ec], i = im[ec];
            re[ec]       = r * sr + i * si;
            im[ec]       = i * sr - r * si;
          }
      }
    // U on every pair bit: (y+, y-) -> (y+ + y-, i (y+ - y-))
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int e = 0; e < NE; ++e)
        if (!((e >> b) & 1))
          {
            const int e1 = e | (1 << b);
            const T   r0 = re[e], i0 = im[e], r1 = re[e1], i1 = im[e1];
            re[e]        = r0 + r1;
            im[e]        = i0 + i1;
            re[e1]       = -(i0 - i1);
            im[e1]       = r0 - r1;
          }
#pragma unroll
    for (int e = 0; e < NE; ++e)
      {
        int a = base;
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if ((e >> b) & 1)
            a += delta[b];
        data[a] = re[e];
      }
  }

  // step 3 for the tasks with NB paired factors of the CPB cells of a
  // workgroup (task k of cell lc on thread k * CPB + lc, strided by the
  // workgroup size); s_cell[lc] = (|a_0|, |a_1|, |a_2|, det J^-1)
  template <int NB, int dim, int n, int NS, int CPB, typename Tc>
  __device__ __forceinline__ void
  fdm_pointwise_group(const FdmDev<Tc> &F, Tc *s_data,
                      const Tc (*s_cell)[4], int64_t cell0, int64_t n_cells,
                      Tc idt)
  {
    constexpr int STR = Lay<dim, n>::stride;
    const int     t0  = F.task_start[NB];
    const int     nt  = (F.task_start[NB + 1] - t0) * CPB;
    for (int t = threadIdx.x; t < nt; t += blockDim.x)
      {
        const int k  = t / CPB;
        const int lc = t - k * CPB;
        if (cell0 + lc >= n_cells)
          continue;
        // one level of independent loads: the task record
        const FdmTask<Tc> &tk   = F.tasks[t0 + k];
        const int          code = tk.code;
        const Tc m[4] = {idt, s_cell[lc][0], s_cell[lc][1], s_cell[lc][2]};
        // factor 0: stage, 1..dim: directions
        int base     = (code & 3) * STR;
        Tc  lr       = tk.re[0] * m[0];
        int delta[4] = {0, 0, 0, 0};
        Tc  li[4]    = {0, 0, 0, 0};
        int nb       = 0;
        if (code & 4)
          {
            delta[0] = STR;
            li[0]    = tk.im[0] * m[0];
            nb       = 1;
          }
        int stride = 1;
#pragma unroll
        for (int d = 0; d < dim; ++d)
          {
            const int f = code >> (3 + 4 * d);
            base += (f & 7) * stride;
            lr += tk.re[1 + d] * m[1 + d];
            if (f & 8)
              {
                // nb < NB <= 4 here
                delta[nb] = stride;
                li[nb]    = tk.im[1 + d] * m[1 + d];
                ++nb;
              }
            stride *= n;
          }
        fdm_block<NB, Tc>(s_data + lc * NS * STR, base, delta, lr, li,
                          s_cell[lc][3]);
      }
  }

  template <int dim, int n, int NS, typename Tc, typename Tin, typename Tout>
  __global__ void __launch_bounds__(128)
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

    const FdmDev<Tc> &F    = *Fp;
    const int         tid  = threadIdx.x;
    const int         lc   = tid / nl;
    const int         p    = tid - lc * nl;
    const int64_t     cell =
      (int64_t)xcd_block(blockIdx.x, gridDim.x) * CPB + lc;
    const bool        active = (lc < CPB) && (cell < P.n_cells);
    const int         lcs    = lc < CPB ? lc : 0;
    Tc               *data   = s_data + lcs * NS * STR;

    // |speed| per direction and det J^-1 of the workgroup's cells, read by
    // any thread in step 3
    __shared__ Tc s_cell[CPB][4];
    bool rv[3] = {false, false, false};
    if (active)
      {
        for (int d = 0; d < dim; ++d)
          {
            const double a = P.cell_velocity[cell * dim + d] * P.vel_scale;
            rv[d]          = a < 0.;
            if (p == 0)
              s_cell[lc][d] = (Tc)fabs(a);
          }
        if (p == 0)
          s_cell[lc][3] = (Tc)P.cell_invdet[cell];
      }

    // 1-2: load (x mirrored), stage transform, V^-1 M^-1 sweeps x, y, (z)
    // with the y and z mirrors on the LDS reads. Every stage has its own
    // LDS block; a thread rewrites exactly the positions of its own line
    // (the mirrors permute within a line), so all stages share one barrier
    // per direction change.
    {
      Tc r[NS][n];
#pragma unroll
      for (int t = 0; t < NS; ++t)
#pragma unroll
        for (int i = 0; i < n; ++i)
          r[t][i] =
            active ? (Tc)src[t * P.block + cell * nd +
                             L::template off<0>(p, mirror<n>(i, rv[0]))] :
                     (Tc)0;
#pragma unroll
      for (int j = 0; j < NS; ++j)
        {
          Tc w[n], t[n];
#pragma unroll
          for (int i = 0; i < n; ++i)
            {
              Tc s = 0;
#pragma unroll
              for (int k = 0; k < NS; ++k)
                s = fma(F.Tinv[j][k], r[k][i], s);
              w[i] = NS == 1 ? r[0][i] : s;
            }
          mat_apply<n>(F.VinvMinv, w, t);
          lds_write_line<dim, n, 0>(data + j * STR, p, t);
        }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < NS; ++j)
        {
          Tc w[n], t[n];
          lds_read_line_m<dim, n, 1>(data + j * STR, p, rv[1], t);
          mat_apply<n>(F.VinvMinv, t, w);
          lds_write_line<dim, n, 1>(data + j * STR, p, w);
        }
      if constexpr (dim == 3)
        {
          __syncthreads();
#pragma unroll
          for (int j = 0; j < NS; ++j)
            {
              Tc w[n], t[n];
              lds_read_line_m<dim, n, 2>(data + j * STR, p, rv[2], t);
              mat_apply<n>(F.VinvMinv, t, w);
              lds_write_line<dim, n, 2>(data + j * STR, p, w);
            }
        }
      __syncthreads();
    }

    // 3: pointwise inverse on the blocks of the workgroup's cells, one
    // group of block shapes (number of paired factors NB) after the other:
    // the threads of a wave run one shape at a time
    {
      const Tc      idt   = (Tc)P.inv_dt;
      const int64_t cell0 = cell - lc;
      fdm_pointwise_group<0, dim, n, NS, CPB>(F, s_data, s_cell, cell0,
                                              P.n_cells, idt);
      fdm_pointwise_group<1, dim, n, NS, CPB>(F, s_data, s_cell, cell0,
                                              P.n_cells, idt);
      fdm_pointwise_group<2, dim, n, NS, CPB>(F, s_data, s_cell, cell0,
                                              P.n_cells, idt);
      if constexpr (dim == 3 || NS > 1)
        fdm_pointwise_group<3, dim, n, NS, CPB>(F, s_data, s_cell, cell0,
                                                P.n_cells, idt);
      if constexpr (dim == 3 && NS > 1)
        fdm_pointwise_group<4, dim, n, NS, CPB>(F, s_data, s_cell, cell0,
                                                P.n_cells, idt);
      __syncthreads();
    }

    // 4: V sweeps LAST, ..., x with the mirrors on the writes; the x
    // mirror goes into the store
    Tc w[NS][n];
#pragma unroll
    for (int j = 0; j < NS; ++j)
      {
        Tc t[n];
        lds_read_line<dim, n, LAST>(data + j * STR, p, t);
        mat_apply<n>(F.V, t, w[j]);
        lds_write_line_m<dim, n, LAST>(data + j * STR, p, rv[LAST], w[j]);
      }
    __syncthreads();
    if constexpr (dim == 3)
      {
#pragma unroll
        for (int j = 0; j < NS; ++j)
          {
            Tc t[n];
            lds_read_line<dim, n, 1>(data + j * STR, p, t);
            mat_apply<n>(F.V, t, w[j]);
            lds_write_line_m<dim, n, 1>(data + j * STR, p, rv[1], w[j]);
          }
        __syncthreads();
      }
#pragma unroll
    for (int j = 0; j < NS; ++j)
      {
        Tc t[n];
        lds_read_line<dim, n, 0>(data + j * STR, p, t);
        mat_apply<n>(F.V, t, w[j]);
      }

    // 5: back stage transform and store (x-lines, mirrored in x)
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
            dst[t * P.block + cell * nd +
                L::template off<0>(p, mirror<n>(i, rv[0]))] =
              (Tout)(NS == 1 ? w[0][i] : s);
          }
  }
} // namespace dgops
