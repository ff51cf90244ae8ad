// Fused matrix-free DG advection operator for gfx950.
//
// One workgroup handles a batch of cells; inside a cell every thread owns
// one 1D line of (k+1) points (n^(dim-1) threads per cell) and keeps it in
// registers. Contractions along the owned line run in registers; changing
// the line direction goes through LDS (write the line, barrier, read the
// line of the new direction).
//
// The operator is evaluated in quadrature space: with S the nodal-to-Gauss
// interpolation matrix, D the collocation derivative at the Gauss points
// and l0/l1 the Gauss-Lagrange values at the faces,
//   integrate(values v, gradients g, face values f) =
//        S^T (v + sum_d D_d^T g_d + sum_faces l_face (x) f)
// so the inverse mass matrix M^-1 = S^-1 diag(1/JxW) S^-T collapses to
//   M^-1 integrate(...) = S^-1 diag(1/JxW) (v + sum_d D_d^T g_d + ...).
// This is the same operator as deal.II's FEEvaluation::integrate followed
// by CellwiseInverseMassMatrix::apply (advection_solver.cc:495-513 and
// :635-657), computed with 9 one-dimensional sweeps per cell instead of
// 12 + 6 in two passes.
//
// Faces are processed element-centric: every cell computes the numerical
// flux on its 2*dim faces with its own outward normal, reading the
// neighbour's face nodes (Gauss-Lobatto nodes make the trace a plain
// subset of the neighbour's values). The flux for the plus side of a
// reference face equals the minus-side flux with the normal reversed
// (advection_solver.cc:560-575), so the result matches the face-centric
// loop of the reference and every dof is written exactly once.
#pragma once

#include <hip/hip_runtime.h>

#include "analytic.hpp"
#include "dev_basis.hpp"
#include "dg_common.hpp"
#include "line_geometry.hpp"

namespace dgops
{
  template <int dim, int n>
  struct Lay
  {
    static constexpr int nl = dim == 3 ? n * n : n; // lines per cell
    static constexpr int nd = nl * n;               // dofs per cell
    static constexpr int stride = nd | 1;           // LDS stride per cell

    // storage offset of point i on line p of direction dir
    template <int dir>
    __device__ static __forceinline__ int
    off(int p, int i)
    {
      if constexpr (dim == 3)
        {
          if constexpr (dir == 0)
            return i + n * p;
          else if constexpr (dir == 1)
            return (p % n) + n * i + n * n * (p / n);
          else
            return p + n * n * i;
        }
      else
        {
          if constexpr (dir == 0)
            return i + n * p;
          else
            return p + n * i;
        }
    }

    template <int dir>
    __device__ static __forceinline__ void
    idx(int p, int i, int (&q)[3])
    {
      if constexpr (dim == 3)
        {
          if constexpr (dir == 0)
            {
              q[0] = i;
              q[1] = p % n;
              q[2] = p / n;
            }
          else if constexpr (dir == 1)
            {
              q[0] = p % n;
              q[1] = i;
              q[2] = p / n;
            }
          else
            {
              q[0] = p % n;
              q[1] = p / n;
              q[2] = i;
            }
        }
      else
        {
          if constexpr (dir == 0)
            {
              q[0] = i;
              q[1] = p;
            }
          else
            {
              q[0] = p;
              q[1] = i;
            }
          q[2] = 0;
        }
    }
  };

  // Workgroups are handed round-robin to the 8 XCDs of the chip, each with
  // its own L2. This bijection of [0, nb) gives every XCD one contiguous
  // range of blocks, so the neighbour cells of a Morton-ordered mesh (and
  // their face traces and tables) are read through the same L2.
  __device__ __forceinline__ unsigned
  xcd_block(unsigned b, unsigned nb)
  {
    constexpr unsigned nx   = 8;
    const unsigned     full = nb / nx, rem = nb % nx;
    const unsigned     x = b % nx, i = b / nx;
    return x < rem ? x * (full + 1) + i : rem * (full + 1) + (x - rem) * full + i;
  }

  // how a kernel obtains the geometry and transport speed at its points
  enum Geometry : int
  {
    geo_table    = 0, // precomputed w adj(J) a, 1/JxW and face tables
    geo_analytic = 1, // axis-aligned cells, speed evaluated per point
    geo_uniform  = 2  // axis-aligned cells, constant speed: products of
                      // 1D weights and per-mesh constants
  };

  // cells per workgroup: two wavefronts' worth of lines, so that every
  // barrier holds back two waves only and more workgroups share a CU
  // (3D degree 4: 5 cells, 125 of 128 lanes busy)
  constexpr int
  block_cells(int lines_per_cell)
  {
    return lines_per_cell >= 128 ? 1 : 128 / lines_per_cell;
  }

  template <int dim, int n>
  constexpr int
  cells_per_block()
  {
    return block_cells(Lay<dim, n>::nl);
  }

  // Tile shapes of the operator kernels, chosen per kernel variant by the
  // autotuner: shape 0 = the default above (two wavefronts), 1 = at most
  // one wavefront of lines, 2 = at most four wavefronts.
  constexpr int n_tile_shapes = 3;

  constexpr int
  shape_cells(int lines_per_cell, int shape)
  {
    return shape == 1 ? (lines_per_cell >= 64 ? 1 : 64 / lines_per_cell) :
           shape == 2 ? (lines_per_cell >= 256 ? 1 : 256 / lines_per_cell) :
                        block_cells(lines_per_cell);
  }

  // ---- small register contractions ---------------------------------------

  // out[q] = sum_i M[q][i] in[i]
  template <int n, typename T>
  __device__ __forceinline__ void
  mat_apply(const T (&M)[max_n][max_n], const T (&in)[n], T (&out)[n])
  {
#pragma unroll
    for (int q = 0; q < n; ++q)
      {
        T s = M[q][0] * in[0];
#pragma unroll
        for (int i = 1; i < n; ++i)
          s = fma(M[q][i], in[i], s);
        out[q] = s;
      }
  }

  // out[i] = sum_q M[q][i] in[q]
  template <int n, typename T>
  __device__ __forceinline__ void
  mat_apply_t(const T (&M)[max_n][max_n], const T (&in)[n], T (&out)[n])
  {
#pragma unroll
    for (int i = 0; i < n; ++i)
      {
        T s = M[0][i] * in[0];
#pragma unroll
        for (int q = 1; q < n; ++q)
          s = fma(M[q][i], in[q], s);
        out[i] = s;
      }
  }

  // out[i] += sum_q M[q][i] in[q]
  template <int n, typename T>
  __device__ __forceinline__ void
  mat_apply_t_add(const T (&M)[max_n][max_n], const T (&in)[n], T (&out)[n])
  {
#pragma unroll
    for (int i = 0; i < n; ++i)
      {
        T s = out[i];
#pragma unroll
        for (int q = 0; q < n; ++q)
          s = fma(M[q][i], in[q], s);
        out[i] = s;
      }
  }

  // Even-odd products with a centrosymmetric (skew = false) or
  // skew-centrosymmetric (skew = true) matrix M (EvenOdd in dev_basis.hpp):
  // out = M in, or M^T in (transpose), added to out (add). With
  // e_i = in_i + in_{n-1-i}, o_i = in_i - in_{n-1-i}, rows q and n-1-q share
  //   se = sum_i E[q][i] e_i + M[q][m] in_m,   so = sum_i O[q][i] o_i
  // as out_q = se + so and out_{n-1-q} = se - so (so - se when skew); the
  // middle row is sum_i M[m][i] e_i + M[m][m] in_m (sum_i M[m][i] o_i when
  // skew). M^T has the same form with E, O transposed (and exchanged when
  // skew) and the middle row and column exchanged.
  template <int n, bool skew, bool transpose, bool add, typename T>
  __device__ __forceinline__ void
  eo_apply(const EvenOdd<T> &M, const T (&in)[n], T (&out)[n])
  {
    constexpr int  h   = n / 2;
    constexpr int  m   = h;
    constexpr bool odd = n % 2 == 1;
    T              e[h > 0 ? h : 1], o[h > 0 ? h : 1];
#pragma unroll
    for (int i = 0; i < h; ++i)
      {
        e[i] = in[i] + in[n - 1 - i];
        o[i] = in[i] - in[n - 1 - i];
      }
    auto E = [&](int q, int i) -> T {
      if constexpr (!transpose)
        return M.e[q][i];
      else if constexpr (!skew)
        return M.e[i][q];
      else
        return M.o[i][q];
    };
    auto O = [&](int q, int i) -> T {
      if constexpr (!transpose)
        return M.o[q][i];
      else if constexpr (!skew)
        return M.o[i][q];
      else
        return M.e[i][q];
    };
    auto col = [&](int q) -> T { return transpose ? M.mr[q] : M.mc[q]; };
    auto row = [&](int i) -> T { return transpose ? M.mc[i] : M.mr[i]; };
#pragma unroll
    for (int q = 0; q < h; ++q)
      {
        T se = odd ? col(q) * in[m] : E(q, 0) * e[0];
#pragma unroll
        for (int i = odd ? 0 : 1; i < h; ++i)
          se = fma(E(q, i), e[i], se);
        T so = O(q, 0) * o[0];
#pragma unroll
        for (int i = 1; i < h; ++i)
          so = fma(O(q, i), o[i], so);
        const T lo = se + so, hi = skew ? so - se : se - so;
        out[q]         = add ? out[q] + lo : lo;
        out[n - 1 - q] = add ? out[n - 1 - q] + hi : hi;
      }
    if constexpr (odd)
      {
        T s;
        if constexpr (skew)
          {
            s = h > 0 ? row(0) * o[0] : (T)0;
#pragma unroll
            for (int i = 1; i < h; ++i)
              s = fma(row(i), o[i], s);
          }
        else
          {
            s = M.mm * in[m];
#pragma unroll
            for (int i = 0; i < h; ++i)
              s = fma(row(i), e[i], s);
          }
        out[m] = add ? out[m] + s : s;
      }
  }

  template <int dim, int n, int dir, typename T>
  __device__ __forceinline__ void
  lds_write_line(T *buf, int p, const T (&v)[n])
  {
#pragma unroll
    for (int i = 0; i < n; ++i)
      buf[Lay<dim, n>::template off<dir>(p, i)] = v[i];
  }

  template <int dim, int n, int dir, typename T>
  __device__ __forceinline__ void
  lds_read_line(const T *buf, int p, T (&v)[n])
  {
#pragma unroll
    for (int i = 0; i < n; ++i)
      v[i] = buf[Lay<dim, n>::template off<dir>(p, i)];
  }

  // row `row` of a cell's geometry table [dim + 1][nd] along line p of the
  // direction the row is stored in (the direction it is read along: row d
  // along d, row dim along x), point i at i * nl + p, so that the lanes of
  // a cell read contiguous memory
  template <int dim, int n, typename T>
  __device__ __forceinline__ void
  table_line(const T *tab, int64_t cell, int row, int p, bool active,
             T (&v)[n])
  {
    constexpr int nd   = Lay<dim, n>::nd;
    constexpr int nl   = Lay<dim, n>::nl;
    const T      *base = tab + (cell * (dim + 1) + row) * nd;
#pragma unroll
    for (int i = 0; i < n; ++i)
      v[i] = active ? base[i * nl + p] : (T)0;
  }

  // ---- geometry ------------------------------------------------------------

  // Per quadrature point data of the cell term: cv[d] = w * (adj(J) a)_d
  // (the reference-space transport speed times JxW, which is what
  // advection_solver_implicit_real.cc:1126-1128 and :998-1007 fold
  // together) and JxW = w det(J), on a multilinear cell (the geometry
  // table builder).
  template <int dim>
  __device__ __forceinline__ void
  cell_point_data(const OpParams &P,
                  const double   *X,
                  const double   *xq,
                  const double   *wq,
                  const int (&q)[3],
                  double (&cv)[3],
                  double &JxW)
  {
    double w = wq[q[0]] * wq[q[1]];
    if (dim == 3)
      w *= wq[q[2]];
    double       x[3], a[3] = {0, 0, 0};
    const double xi[3] = {xq[q[0]], xq[q[1]], dim == 3 ? xq[q[2]] : 0.};
    double       J[3][3], adj[3][3];
    multilinear_map<dim>(X, xi, x, J);
    velocity_field<dim>(P.vel_kind, P.vel_const, x, a);
    const double det = adjugate<dim>(J, adj);
    for (int d = 0; d < dim; ++d)
      {
        double s = 0;
        for (int e = 0; e < dim; ++e)
          s += adj[d][e] * a[e];
        cv[d] = w * s;
      }
    JxW = w * det;
  }

  // Face point data: an = a . n dA (outward normal of this cell, including
  // the face quadrature weight) and the physical point x, on a multilinear
  // cell (the face and boundary table builders).
  template <int dim>
  __device__ __forceinline__ void
  face_point_data(const OpParams &P,
                  const double   *X,
                  const double   *xq,
                  const double   *wq,
                  int             dir,
                  int             side,
                  const int (&q)[3],
                  double &an,
                  double (&x)[3])
  {
    double w = 1.;
    for (int e = 0; e < dim; ++e)
      if (e != dir)
        w *= wq[q[e]];
    double xi[3];
    for (int e = 0; e < 3; ++e)
      xi[e] = (e < dim) ? xq[q[e]] : 0.;
    xi[dir] = side ? 1. : 0.;
    double a[3] = {0, 0, 0};
    double J[3][3], adj[3][3];
    multilinear_map<dim>(X, xi, x, J);
    velocity_field<dim>(P.vel_kind, P.vel_const, x, a);
    adjugate<dim>(J, adj);
    double s = 0;
    for (int e = 0; e < dim; ++e)
      s += adj[dir][e] * a[e];
    an = (side ? 1. : -1.) * w * s;
  }


  // ---- face terms ------------------------------------------------------------

  // Numerical flux on the two faces normal to `dir`, at the face point owned
  // by this thread in the dir-line layout; adds sign * l_side (x) flux to the
  // line accumulator. Positive-operator convention of
  // advection_solver_implicit_real.cc:1032-1058:
  //   F = 1/2 [(u- + u+) an + alpha |an| (u- - u+)] - skew an u-
  // with an = a . n dA on this cell's outward normal.
  template <int dim, int n, int dir, typename T, bool diff, bool local = false>
  __device__ __forceinline__ void
  face_terms(const OpParams &P,
             const DevBasis<double> *Bd,
             const DevBasis<T>      &B,
             int64_t                 cell,
             int                     p,
             bool                    active,
             const int (&nb)[2 * dim],
             const T (&uplus)[2 * dim],
             const T (&dplus)[2 * dim],
             const T (&uq)[n],
             T (&acc)[n],
             T    vf,
             T (&fan)[2 * dim])
  {
    int q[3];
    Lay<dim, n>::template idx<dir>(p, 0, q);
#pragma unroll
    for (int s = 0; s < 2; ++s)
      {
        const int f = 2 * dir + s;
        const T(&l)[max_n] = s ? B.l1 : B.l0;
        T um = l[0] * uq[0];
#pragma unroll
        for (int i = 1; i < n; ++i)
          um = fma(l[i], uq[i], um);
        T up = uplus[f];
        if (!local && active && nb[f] < 0)
          // Dirichlet data, evaluated before the launch
          // (boundary_value_kernel)
          up = (T)P.bnd_value[(int64_t)P.bnd_index[cell * 2 * dim + f] *
                                Lay<dim, n>::nl +
                              p];
        const T an = fan[f] * vf;
        const T flux =
          (T)0.5 * ((um + up) * an + (T)P.alpha * fabs(an) * (um - up)) -
          (T)P.skew * an * um;
        const T sf = (T)P.sign * flux;
#pragma unroll
        for (int i = 0; i < n; ++i)
          acc[i] = fma(l[i], sf, acc[i]);
        if constexpr (diff)
          {
            // SIPG (advection_diffusion_solver.cc:569-613) on this cell's
            // outward normal, axis-aligned cells: with
            //   jump = -nu/2 (u- - u+), avg = -nu/2 (d_n u- + d_n u+)
            // the value term is -(jump sigma - avg) and the normal
            // derivative of the test function is weighted with jump
            const T(&dl)[max_n] = s ? B.dl1 : B.dl0;
            const T hd  = (T)P.h[dir];
            const T sgn = s ? (T)1 : (T)-1;
            T       dm  = dl[0] * uq[0];
#pragma unroll
            for (int i = 1; i < n; ++i)
              dm = fma(dl[i], uq[i], dm);
            const T dnm = sgn * dm / hd, dnp = sgn * dplus[f] / hd;
            T       w   = (T)1;
            for (int e = 0; e < dim; ++e)
              if (e != dir)
                w *= (T)(Bd->w[q[e]] * P.h[e]);
            const T nu    = (T)P.viscosity;
            const T sigma = (T)P.penalty / hd;
            const T jump  = (T)-0.5 * nu * (um - up);
            const T avg   = (T)-0.5 * nu * (dnm + dnp);
            const T fv    = (T)P.sign * w * (avg - jump * sigma);
            const T fd    = (T)P.sign * w * jump * sgn / hd;
#pragma unroll
            for (int i = 0; i < n; ++i)
              acc[i] = fma(l[i], fv, fma(dl[i], fd, acc[i]));
          }
      }
  }

  // cell diffusion term along direction dir on axis-aligned cells:
  // acc += D^T (sign nu |K|/h_d^2 w_q D u_q)  (int grad v . nu grad u)
  template <int dim, int n, int dir, typename T>
  __device__ __forceinline__ void
  diffusion_cell_term(const OpParams &P, const DevBasis<double> *Bd,
                      const DevBasis<T> &B, int p, const T (&uq)[n],
                      T (&acc)[n])
  {
    double vol = P.h[0] * P.h[1];
    if (dim == 3)
      vol *= P.h[2];
    const T c0 = (T)(P.sign * P.viscosity * vol / (P.h[dir] * P.h[dir]));
    T       du[n], g[n];
    eo_apply<n, true, false, false>(B.eo_D, uq, du);
#pragma unroll
    for (int i = 0; i < n; ++i)
      {
        int q[3];
        Lay<dim, n>::template idx<dir>(p, i, q);
        T w = (T)(Bd->w[q[0]] * Bd->w[q[1]]);
        if (dim == 3)
          w *= (T)Bd->w[q[2]];
        g[i] = c0 * w * du[i];
      }
    eo_apply<n, true, true, true>(B.eo_D, g, acc);
  }

  // ---- the kernel ------------------------------------------------------------

  enum Finalize : int
  {
    fin_inverse_mass = 0, // S^-1 diag(1/JxW): explicit right-hand side
    fin_integrate    = 1, // S^T: operator action (implicit vmult)
    fin_local        = 2  // cell-local operator in the Gauss-point
                          // collocation basis: input and output at the
                          // Gauss points, neighbour traces zero
                          // (CellwiseOperator::vmult,
                          // advection_solver_irk_real.cc:545-711)
  };

  enum Epilogue : int
  {
    epi_store = 0, // dst = scale * y
    epi_lsrk  = 1  // k = y; sol_out = sol_in + b k; ri_out = sol_in + a k
  };

  template <typename T>
  struct OpBuffers
  {
    const T *src; // [NC][comp_stride]
    T       *dst; // [NC][comp_stride]
    int64_t  comp_stride;
    T        scale;
    // LSRK epilogue (NC == 1)
    const T *sol_in;
    T       *sol_out;
    T       *ri_out;
    T        b_dt, a_dt;
    int      last;
  };

  // y_c = fin( sign * L_c u_c + sum_b mass[c][b] M u_b ), c < NC, with
  // L_c the advection operator at velocity vel_factor[c] * a(x).
  // With `diff` (axis-aligned cells) L also contains the SIPG diffusion
  // operator of advection_diffusion_solver.cc:531-613.
  // The operator on the cells of workgroup `block` (one batch of CPB
  // cells); P and io live in kernel arguments or, for the
  // batched launch, in global memory (uniform per workgroup either way).
  template <int dim, int n, int NC, typename T, int geo, bool skew,
            bool diff, int fin, int epi, int CPB>
  __device__ __forceinline__ void
  operator_body(const OpParams                       &P,
                const DevBasis<T> *__restrict__      Bp,
                const DevBasis<double> *__restrict__ Bd,
                const OpBuffers<T>                   &io,
                unsigned                              block)
  {
    static_assert(!diff || geo != geo_table, "SIPG needs axis-aligned cells");
    static_assert(fin != fin_local || (!diff && epi == epi_store),
                  "the cell-local operator is built for advection");
    using L            = Lay<dim, n>;
    constexpr int nl   = L::nl;
    constexpr int nd   = L::nd;
    constexpr int STR  = L::stride;
    constexpr int FSTR = (diff ? 2 : 1) * 2 * dim * nl;
    constexpr int NV   = 1 << dim;
    constexpr int LAST = dim - 1;
    constexpr int NX   = geo == geo_table ? NV * dim : dim;

    // analytic (axis-aligned) geometry evaluates the transport speed at
    // LAST-line points and hands the other directions over through LDS;
    // tabulated and uniform geometry produce every direction in its own
    // line layout
    constexpr bool AN  = geo == geo_analytic;
    constexpr bool UNI = geo == geo_uniform;
    __shared__ T      s_u[CPB * STR];
    __shared__ T      s_acc[CPB * STR];
    __shared__ T      s_g0[AN ? CPB * STR : 1];
    __shared__ T      s_g1[(AN && dim == 3) ? CPB * STR : 1];
    __shared__ T      s_sk0[(AN && skew) ? CPB * STR : 1];
    __shared__ T      s_sk1[(AN && skew && dim == 3) ? CPB * STR : 1];
    __shared__ T      s_face[dim == 3 ? CPB * FSTR : 1];
    __shared__ T      s_fan[AN ? CPB * 2 * (dim - 1) * nl : 1];
    __shared__ double s_X[CPB * NX];
    // several components (IRK stages): their values at the quadrature
    // points wait in LDS between the value phase and the operator phase
    // (LAST-line layout: every thread writes and reads its own line only)
    // instead of holding NC lines in registers through the operator phase
    constexpr bool STAGE_LDS = NC > 1;
    __shared__ T   s_stage[STAGE_LDS ? NC * CPB * STR : 1];
    // several components, 3D advection: the neighbour traces of component
    // c are interpolated at the start of its operator phase instead of in
    // the value phase, so that one component's traces are live at a time
    // (NC sets of traces held through the operator phase cost one
    // occupancy step)
    constexpr bool TRACE_LATE =
      STAGE_LDS && dim == 3 && !diff && fin != fin_local;
    const T *tab_c = static_cast<const T *>(P.tab_cell);
    const T *tab_f = static_cast<const T *>(P.tab_face);

    const DevBasis<T> &B    = *Bp;
    const int          tid  = threadIdx.x;
    const int          lc   = tid / nl;
    const int          p    = tid - lc * nl;
    const int64_t      cell = (int64_t)block * CPB + lc;
    const bool         active = (lc < CPB) && (cell < P.n_cells);
    const int          lcs    = lc < CPB ? lc : 0;

    T      *u_l   = s_u + lcs * STR;
    T      *acc_l = s_acc + lcs * STR;
    T      *g0_l  = AN ? s_g0 + lcs * STR : s_g0;
    T      *g1_l  = (AN && dim == 3) ? s_g1 + lcs * STR : s_g1;
    T      *sk0_l = (AN && skew) ? s_sk0 + lcs * STR : s_sk0;
    T      *sk1_l = (AN && skew && dim == 3) ? s_sk1 + lcs * STR : s_sk1;
    T      *f_l   = dim == 3 ? s_face + lcs * FSTR : s_face;
    T      *fan_l = AN ? s_fan + lcs * 2 * (dim - 1) * nl : s_fan;
    double *X     = s_X + lcs * NX;
    auto    stage_l = [&](int c) -> T * {
      return s_stage + (STAGE_LDS ? (c * CPB + lcs) * STR : 0);
    };

    // the cell's own values of the first component are requested first:
    // the x sweep then overlaps the neighbour-index -> trace load chain
    // (values at the Gauss points of the collocation basis are read along
    // LAST-direction lines, where the operator starts)
    constexpr int in_dir = fin == fin_local ? LAST : 0;
    auto          load_own = [&](const T *src, T(&r)[n]) {
#pragma unroll
      for (int i = 0; i < n; ++i)
        r[i] =
          active ? src[cell * nd + L::template off<in_dir>(p, i)] : (T)0;
    };
    // waves that are issuing global loads (the cell's values and traces,
    // the tables) or stores get issue priority, so that requests leave
    // early while other waves compute
    __builtin_amdgcn_s_setprio(3);
    T r0[n];
    load_own(io.src, r0);
    // cell vertices: analytic geometry, and JxW of the mass term with
    // tabulated geometry
    constexpr bool needX = AN || (geo == geo_table && fin != fin_inverse_mass);
    if constexpr (needX)
      {
        if (active)
          for (int e = p; e < NX; e += nl)
            X[e] = P.cell_vertices[cell * NV * dim + e];
      }
    int nb[2 * dim];
    if (dim == 3 && P.morton_bits >= 0)
      {
        // one periodic coarse cell: the neighbours are Morton increments
        // of the cell index (dilated integers, carries confined to one
        // coordinate's bits), so the trace loads below need no index load
        // first
        const unsigned full = (1u << (3 * P.morton_bits)) - 1u;
        const unsigned m    = (unsigned)cell;
#pragma unroll
        for (int d = 0; d < dim; ++d)
          {
            const unsigned M = (0x09249249u << d) & full;
            nb[2 * d] = active ? (int)((((m & M) - 1u) & M) | (m & ~M)) : -1;
            nb[2 * d + 1] =
              active ? (int)((((m | ~M) + 1u) & M) | (m & ~M)) : -1;
          }
      }
    else
#pragma unroll
      for (int f = 0; f < 2 * dim; ++f)
        nb[f] = active ? P.neighbors[cell * 2 * dim + f] : -1;
    if constexpr (needX)
      __syncthreads();

    // tabulated geometry: LAST-direction speed on this thread's line, the
    // y and x speed rows, and a . n dA on the faces its lines end on (lower
    // face of this cell, upper face = lower face of the upper neighbour,
    // boundary faces apart). Requested after the value phase: holding them
    // across its barriers costs an occupancy step, which loses more than
    // the overlap gains, with one component and with several.
    T cv[3][n], jxw[n], cvy[n], cvx[n];
    T fan[2 * dim];
    auto load_tables = [&]() {
        table_line<dim, n>(tab_c, cell, LAST, p, active, cv[LAST]);
#pragma unroll
        for (int d = 0; d < dim; ++d)
          {
            fan[2 * d] =
              active ? -tab_f[(cell * dim + d) * nl + p] : (T)0;
            T up = 0;
            if (active)
              {
                const int nbu = nb[2 * d + 1];
                if (nbu >= 0)
                  up = tab_f[((int64_t)nbu * dim + d) * nl + p];
                else
                  up = (T)P.tab_bnd[((int64_t)P.bnd_index[cell * 2 * dim +
                                                          2 * d + 1] *
                                       nl +
                                     p) *
                                    (1 + dim)];
              }
            fan[2 * d + 1] = up;
            if (active && nb[2 * d] < 0)
              fan[2 * d] =
                (T)P.tab_bnd[((int64_t)P.bnd_index[cell * 2 * dim + 2 * d] *
                                nl +
                              p) *
                             (1 + dim)];
          }
    };
    auto load_speed_rows = [&]() {
      if constexpr (dim == 3)
        table_line<dim, n>(tab_c, cell, 1, p, active, cvy);
      table_line<dim, n>(tab_c, cell, 0, p, active, cvx);
    };

    // ---- per component: neighbour traces and values at quadrature points
    constexpr int NUP = TRACE_LATE ? 1 : NC;
    T             uplus[NUP][2 * dim];
    T             dplus[NUP][2 * dim];
    // values at the quadrature points on this thread's LAST line: in
    // registers for one component, in s_stage for several
    T    uqs[STAGE_LDS ? 1 : NC][n];
    auto keep_uq = [&](int c, const T(&v)[n]) {
      if constexpr (STAGE_LDS)
        lds_write_line<dim, n, LAST>(stage_l(c), p, v);
      else
#pragma unroll
        for (int i = 0; i < n; ++i)
          uqs[0][i] = v[i];
    };
    auto get_uq = [&](int c, T(&v)[n]) {
      if constexpr (STAGE_LDS)
        lds_read_line<dim, n, LAST>(stage_l(c), p, v);
      else
#pragma unroll
        for (int i = 0; i < n; ++i)
          v[i] = uqs[0][i];
    };
#pragma unroll
    for (int c = 0; c < NC; ++c)
      {
        const T *src = io.src + c * io.comp_stride;
        T        r[n], t[n];
        if (c == 0)
#pragma unroll
          for (int i = 0; i < n; ++i)
            r[i] = r0[i];
        else
          load_own(src, r);
        if constexpr (fin == fin_local)
          {
            // already at the Gauss points; the neighbours do not couple
            keep_uq(c, r);
#pragma unroll
            for (int f = 0; f < 2 * dim; ++f)
              uplus[c][f] = dplus[c][f] = 0;
            continue;
          }
        // neighbours' face nodes for LDS staging (thread (i, j) loads node
        // (t1 = i, t2 = j) of every face, one load per face); stored after
        // the x sweep so that the sweep does not wait for them
        T tv[6];
        if constexpr (dim == 3 && !diff && !TRACE_LATE)
          {
            const int i0 = p % n, j0 = p / n;
#pragma unroll
            for (int f = 0; f < 6; ++f)
              {
                const int d = f / 2, s = f % 2;
                const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
                tv[f]        = 0;
                if (active && nb[f] >= 0)
                  {
                    int ijk[3];
                    ijk[d]  = s ? 0 : n - 1;
                    ijk[t1] = i0;
                    ijk[t2] = j0;
                    tv[f]   = src[(int64_t)nb[f] * nd + ijk[0] + n * ijk[1] +
                                n * n * ijk[2]];
                  }
              }
          }
        // own values, sweep along x
        __builtin_amdgcn_s_setprio(0);
        eo_apply<n, false, false, false>(B.eo_S, r, t);
        lds_write_line<dim, n, 0>(acc_l, p, t);
        if constexpr (dim == 3 && !diff)
          {
            if constexpr (!TRACE_LATE)
              {
                const int i0 = p % n, j0 = p / n;
#pragma unroll
                for (int f = 0; f < 6; ++f)
                  f_l[f * nl + j0 * n + i0] = tv[f];
              }
          }
        else if constexpr (dim == 3)
          {
            // step 1: thread (aq, j) interpolates the neighbour's face nodes
            // along the first tangential direction
            const int aq = p % n, j = p / n;
#pragma unroll
            for (int f = 0; f < 6; ++f)
              {
                const int d = f / 2, s = f % 2;
                const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
                T         v  = 0;
                if (active && nb[f] >= 0)
                  {
                    const T  *nsrc = src + (int64_t)nb[f] * nd;
                    const int nn   = s ? 0 : n - 1;
#pragma unroll
                    for (int i = 0; i < n; ++i)
                      {
                        int ijk[3];
                        ijk[d]  = nn;
                        ijk[t1] = i;
                        ijk[t2] = j;
                        v = fma(B.S[aq][i],
                                nsrc[ijk[0] + n * ijk[1] + n * n * ijk[2]], v);
                      }
                  }
                f_l[f * nl + j * n + aq] = v;
                if constexpr (diff)
                  {
                    // reference-space derivative of the neighbour normal
                    // to the face, at its face (nodal derivative at 1 or 0)
                    const T(&dp)[max_n] = s ? B.dphi0 : B.dphi1;
                    T         dv        = 0;
                    if (active && nb[f] >= 0)
                      {
                        const T *nsrc = src + (int64_t)nb[f] * nd;
#pragma unroll
                        for (int i = 0; i < n; ++i)
                          {
                            T dd = 0;
#pragma unroll
                            for (int l = 0; l < n; ++l)
                              {
                                int ijk[3];
                                ijk[d]  = l;
                                ijk[t1] = i;
                                ijk[t2] = j;
                                dd      = fma(dp[l],
                                         nsrc[ijk[0] + n * ijk[1] +
                                              n * n * ijk[2]],
                                         dd);
                              }
                            dv = fma(B.S[aq][i], dd, dv);
                          }
                      }
                    f_l[(6 + f) * nl + j * n + aq] = dv;
                  }
              }
          }
        __syncthreads();
        T fv[6];
        if constexpr (dim == 3 && !diff)
          {
            if constexpr (!TRACE_LATE)
            {
            // first tangential interpolation of the staged nodes: thread
            // (aq, j); written back after the next barrier
            const int aq = p % n, j = p / n;
#pragma unroll
            for (int f = 0; f < 6; ++f)
              {
                T v = 0;
#pragma unroll
                for (int i = 0; i < n; ++i)
                  v = fma(B.S[aq][i], f_l[f * nl + j * n + i], v);
                fv[f] = v;
              }
            }
          }
        else if constexpr (dim == 3)
          {
            // step 2 of the face interpolation: thread (aq, bq)
            const int aq = p % n, bq = p / n;
#pragma unroll
            for (int f = 0; f < 6; ++f)
              {
                T v = 0;
#pragma unroll
                for (int jj = 0; jj < n; ++jj)
                  v = fma(B.S[bq][jj], f_l[f * nl + jj * n + aq], v);
                uplus[c][f] = v;
                if constexpr (diff)
                  {
                    T dv = 0;
#pragma unroll
                    for (int jj = 0; jj < n; ++jj)
                      dv = fma(B.S[bq][jj], f_l[(6 + f) * nl + jj * n + aq],
                               dv);
                    dplus[c][f] = dv;
                  }
              }
          }
        else
          {
#pragma unroll
            for (int f = 0; f < 4; ++f)
              {
                const int d = f / 2, s = f % 2;
                T         v = 0, dv = 0;
                if (active && nb[f] >= 0)
                  {
                    const T  *nsrc = src + (int64_t)nb[f] * nd;
                    const int nn   = s ? 0 : n - 1;
                    const T(&dp)[max_n] = s ? B.dphi0 : B.dphi1;
#pragma unroll
                    for (int i = 0; i < n; ++i)
                      {
                        const int ij0 = d == 0 ? nn : i;
                        const int ij1 = d == 0 ? i : nn;
                        v = fma(B.S[p][i], nsrc[ij0 + n * ij1], v);
                        if constexpr (diff)
                          {
                            T dd = 0;
#pragma unroll
                            for (int l = 0; l < n; ++l)
                              dd = fma(dp[l],
                                       nsrc[d == 0 ? l + n * i : i + n * l],
                                       dd);
                            dv = fma(B.S[p][i], dd, dv);
                          }
                      }
                  }
                uplus[c][f] = v;
                dplus[c][f] = dv;
              }
          }
        lds_read_line<dim, n, 1>(acc_l, p, t);
        eo_apply<n, false, false, false>(B.eo_S, t, r);
        if constexpr (dim == 3)
          {
            __syncthreads();
            if constexpr (!diff && !TRACE_LATE)
              {
                const int aq = p % n, j = p / n;
#pragma unroll
                for (int f = 0; f < 6; ++f)
                  f_l[f * nl + j * n + aq] = fv[f];
              }
            lds_write_line<dim, n, 1>(acc_l, p, r);
            __syncthreads();
            if constexpr (!diff && !TRACE_LATE)
              {
                // second tangential interpolation: thread (aq, bq)
                const int aq = p % n, bq = p / n;
#pragma unroll
                for (int f = 0; f < 6; ++f)
                  {
                    T v = 0;
#pragma unroll
                    for (int jj = 0; jj < n; ++jj)
                      v = fma(B.S[bq][jj], f_l[f * nl + jj * n + aq], v);
                    uplus[c][f] = v;
                  }
              }
            lds_read_line<dim, n, 2>(acc_l, p, t);
            eo_apply<n, false, false, false>(B.eo_S, t, r);
          }
        keep_uq(c, r);
        __syncthreads();
      }

    // ---- geometry and velocity: this thread's LAST-line points and the
    // faces its line ends on; faces normal to the other directions are
    // evaluated along LAST-direction lines by up to one task per thread and
    // handed over through LDS
    // uniform geometry: w adj(J) a at point (i, line p of direction d) is
    // w_i wl_d(p) |K| / h_d a_d, with wl_d(p) the product of the 1D
    // weights at the line's other coordinates; a . n dA on the faces of
    // that line is -+ wl_d(p) |K| / h_d a_d
    double vol = 0, cdir[3] = {0, 0, 0}, wl[3] = {1, 1, 1};
    if constexpr (UNI)
      {
        vol = P.h[0] * P.h[1] * (dim == 3 ? P.h[2] : 1.);
#pragma unroll
        for (int d = 0; d < dim; ++d)
          cdir[d] = vol / P.h[d] * P.vel_const[d];
        int q[3];
#pragma unroll
        for (int d = 0; d < dim; ++d)
          {
            if (d == 0)
              L::template idx<0>(p, 0, q);
            else if (d == 1)
              L::template idx<(dim > 1 ? 1 : 0)>(p, 0, q);
            else
              L::template idx<LAST>(p, 0, q);
#pragma unroll
            for (int e = 0; e < dim; ++e)
              if (e != d)
                wl[d] *= Bd->w[q[e]];
          }
      }
    if constexpr (UNI)
      {
        const double cl = wl[LAST] * cdir[LAST];
#pragma unroll
        for (int i = 0; i < n; ++i)
          {
            cv[LAST][i] = (T)(Bd->w[i] * cl);
            jxw[i] = fin != fin_inverse_mass ? (T)(Bd->w[i] * wl[LAST] * vol)
                                          : (T)1;
          }
#pragma unroll
        for (int d = 0; d < dim; ++d)
          {
            fan[2 * d]     = (T)(-wl[d] * cdir[d]);
            fan[2 * d + 1] = (T)(wl[d] * cdir[d]);
          }
      }
    else if constexpr (!AN)
      {
        __builtin_amdgcn_s_setprio(3);
        load_tables();
        __builtin_amdgcn_s_setprio(0);
        // JxW of the mass term on this thread's line
        if (fin != fin_inverse_mass && active && P.has_mass)
          line_jxw<dim, n, LAST, T, L>(X, Bd->xq, Bd->w, p, jxw);
        else
#pragma unroll
          for (int i = 0; i < n; ++i)
            jxw[i] = 1;
      }
    else
    {
      int q[3];
      L::template idx<LAST>(p, 0, q);
      T fl[2] = {0, 0};
      if (active)
        line_cell_data<dim, n, LAST, true, T>(P, X, Bd, q, cv, jxw, fl);
      else
        {
#pragma unroll
          for (int i = 0; i < n; ++i)
            {
              cv[0][i] = cv[1][i] = cv[2][i] = 0;
              jxw[i]                         = 1;
            }
        }
      fan[2 * LAST]     = fl[0];
      fan[2 * LAST + 1] = fl[1];
      constexpr int ntask = dim == 3 ? 4 * n : 2;
      if (active)
        for (int task = p; task < ntask; task += nl)
          {
            const int d = dim == 3 ? task / (2 * n) : 0;
            const int s = dim == 3 ? (task / n) % 2 : task;
            const int j = dim == 3 ? task % n : 0;
            face_line_data<dim, n, LAST, true, T>(
              P, X, Bd, d, s, j, fan_l + (2 * d + s) * nl + j * n);
          }
      __syncthreads();
      const int fidx = dim == 3 ? (p % n) * n + p / n : p;
#pragma unroll
      for (int f = 0; f < 2 * LAST; ++f)
        fan[f] = fan_l[f * nl + fidx];
    }

    // speed rows of the y and x phases, read once for all components
    if constexpr (geo == geo_table)
      load_speed_rows();
    // tabulated inverse mass: the 1/JxW row is requested after the LAST
    // phase (one component; several components read it per component
    // after their x phase)
    constexpr bool early_ij = geo == geo_table && fin == fin_inverse_mass;
    T ij_row[early_ij ? n : 1];
    if constexpr (UNI)
      {
#pragma unroll
        for (int i = 0; i < n; ++i)
          {
            cvy[i] = (T)(Bd->w[i] * wl[dim == 3 ? 1 : 0] * cdir[1]);
            cvx[i] = (T)(Bd->w[i] * wl[0] * cdir[0]);
          }
      }
#pragma unroll
    for (int c = 0; c < NC; ++c)
      {
        if constexpr (TRACE_LATE)
          {
            // neighbour traces of component c: thread (i0, j0) stages node
            // (t1 = i0, t2 = j0) of each face, then interpolates along t1
            // (thread (aq = i0, j0)) and along t2 (thread (aq = i0, bq = j0))
            const T  *src = io.src + c * io.comp_stride;
            const int i0 = p % n, j0 = p / n;
            T         tv[6];
            __builtin_amdgcn_s_setprio(3);
#pragma unroll
            for (int f = 0; f < 6; ++f)
              {
                const int d = f / 2, s = f % 2;
                const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
                tv[f]        = 0;
                if (active && nb[f] >= 0)
                  {
                    int ijk[3];
                    ijk[d]  = s ? 0 : n - 1;
                    ijk[t1] = i0;
                    ijk[t2] = j0;
                    tv[f]   = src[(int64_t)nb[f] * nd + ijk[0] + n * ijk[1] +
                                n * n * ijk[2]];
                  }
              }
            __builtin_amdgcn_s_setprio(0);
#pragma unroll
            for (int f = 0; f < 6; ++f)
              f_l[f * nl + j0 * n + i0] = tv[f];
            __syncthreads();
#pragma unroll
            for (int f = 0; f < 6; ++f)
              {
                T v = 0;
#pragma unroll
                for (int i = 0; i < n; ++i)
                  v = fma(B.S[i0][i], f_l[f * nl + j0 * n + i], v);
                tv[f] = v;
              }
            __syncthreads();
#pragma unroll
            for (int f = 0; f < 6; ++f)
              f_l[f * nl + j0 * n + i0] = tv[f];
            __syncthreads();
#pragma unroll
            for (int f = 0; f < 6; ++f)
              {
                T v = 0;
#pragma unroll
                for (int jj = 0; jj < n; ++jj)
                  v = fma(B.S[j0][jj], f_l[f * nl + jj * n + i0], v);
                uplus[0][f] = v;
              }
          }
        const T vf    = (T)P.vel_factor[c];
        const T fcell = (T)(-(1. - P.skew) * P.sign) * vf;
        const T fs    = (T)(P.skew * P.sign) * vf;
        T       acc[n];
        // ---- LAST direction (z in 3D, y in 2D) ----
        {
#pragma unroll
          for (int i = 0; i < n; ++i)
            acc[i] = 0;
          if (fin != fin_inverse_mass && P.has_mass)
            {
#pragma unroll
              for (int b = 0; b < NC; ++b)
                {
                  const T m = (T)P.mass[c][b];
                  T       ub[n];
                  get_uq(b, ub);
#pragma unroll
                  for (int i = 0; i < n; ++i)
                    acc[i] = fma(m * jxw[i], ub[i], acc[i]);
                }
            }
          T uc[n];
          get_uq(c, uc);
          T g[n];
#pragma unroll
          for (int i = 0; i < n; ++i)
            g[i] = fcell * cv[LAST][i] * uc[i];
          eo_apply<n, true, true, true>(B.eo_D, g, acc);
          if constexpr (skew)
            {
              T du[n];
              eo_apply<n, true, false, false>(B.eo_D, uc, du);
#pragma unroll
              for (int i = 0; i < n; ++i)
                acc[i] = fma(fs * cv[LAST][i], du[i], acc[i]);
            }
          if constexpr (diff)
            diffusion_cell_term<dim, n, LAST, T>(P, Bd, B, p, uc, acc);
          face_terms<dim, n, LAST, T, diff, fin == fin_local>(P, Bd, B, cell, p, active, nb,
                                            uplus[NUP > 1 ? c : 0],
                                            dplus[NUP > 1 ? c : 0], uc, acc,
                                            vf, fan);
          if constexpr (early_ij)
            {
              // the 1/JxW row of the inverse mass, requested here (the
              // LAST speed row is dead) so that it arrives during the y
              // and x phases instead of after them; no extra registers at
              // the kernel's peak (3D degree 4: 107 VGPRs either way)
              __builtin_amdgcn_s_setprio(3);
              table_line<dim, n>(tab_c, cell, dim, p, active, ij_row);
              __builtin_amdgcn_s_setprio(0);
            }
          lds_write_line<dim, n, LAST>(u_l, p, uc);
          lds_write_line<dim, n, LAST>(acc_l, p, acc);
          if constexpr (AN)
            {
#pragma unroll
              for (int i = 0; i < n; ++i)
                g[i] = fcell * cv[0][i] * uc[i];
              lds_write_line<dim, n, LAST>(g0_l, p, g);
              if constexpr (dim == 3)
                {
#pragma unroll
                  for (int i = 0; i < n; ++i)
                    g[i] = fcell * cv[1][i] * uc[i];
                  lds_write_line<dim, n, LAST>(g1_l, p, g);
                }
              if constexpr (skew)
                {
#pragma unroll
                  for (int i = 0; i < n; ++i)
                    g[i] = fs * cv[0][i];
                  lds_write_line<dim, n, LAST>(sk0_l, p, g);
                  if constexpr (dim == 3)
                    {
#pragma unroll
                      for (int i = 0; i < n; ++i)
                        g[i] = fs * cv[1][i];
                      lds_write_line<dim, n, LAST>(sk1_l, p, g);
                    }
                }
            }
          __syncthreads();
        }

        // ---- y direction in 3D ----
        if constexpr (dim == 3)
          {
            T        uq[n], g[n];
            const T(&cvd)[n] = cvy;
            lds_read_line<dim, n, 1>(u_l, p, uq);
            if constexpr (AN)
              lds_read_line<dim, n, 1>(g1_l, p, g);
            else
              {
#pragma unroll
                for (int i = 0; i < n; ++i)
                  g[i] = fcell * cvd[i] * uq[i];
              }
            lds_read_line<dim, n, 1>(acc_l, p, acc);
            eo_apply<n, true, true, true>(B.eo_D, g, acc);
            if constexpr (skew)
              {
                T du[n], sk[n];
                eo_apply<n, true, false, false>(B.eo_D, uq, du);
                if constexpr (AN)
                  lds_read_line<dim, n, 1>(sk1_l, p, sk);
                else
#pragma unroll
                  for (int i = 0; i < n; ++i)
                    sk[i] = fs * cvd[i];
#pragma unroll
                for (int i = 0; i < n; ++i)
                  acc[i] = fma(sk[i], du[i], acc[i]);
              }
            if constexpr (diff)
              diffusion_cell_term<dim, n, 1, T>(P, Bd, B, p, uq, acc);
            face_terms<dim, n, 1, T, diff, fin == fin_local>(P, Bd, B, cell, p, active, nb,
                                           uplus[NUP > 1 ? c : 0],
                                           dplus[NUP > 1 ? c : 0], uq, acc, vf,
                                           fan);
            lds_write_line<dim, n, 1>(acc_l, p, acc);
            __syncthreads();
          }

        // ---- x direction ----
        {
          T        uq[n], g[n];
          const T(&cvd)[n] = cvx;
          lds_read_line<dim, n, 0>(u_l, p, uq);
          if constexpr (AN)
            lds_read_line<dim, n, 0>(g0_l, p, g);
          else
            {
#pragma unroll
              for (int i = 0; i < n; ++i)
                g[i] = fcell * cvd[i] * uq[i];
            }
          lds_read_line<dim, n, 0>(acc_l, p, acc);
          eo_apply<n, true, true, true>(B.eo_D, g, acc);
          if constexpr (skew)
            {
              T du[n], sk[n];
              eo_apply<n, true, false, false>(B.eo_D, uq, du);
              if constexpr (AN)
                lds_read_line<dim, n, 0>(sk0_l, p, sk);
              else
#pragma unroll
                for (int i = 0; i < n; ++i)
                  sk[i] = fs * cvd[i];
#pragma unroll
              for (int i = 0; i < n; ++i)
                acc[i] = fma(sk[i], du[i], acc[i]);
            }
          if constexpr (diff)
            diffusion_cell_term<dim, n, 0, T>(P, Bd, B, p, uq, acc);
          face_terms<dim, n, 0, T, diff, fin == fin_local>(P, Bd, B, cell, p, active, nb,
                                         uplus[NUP > 1 ? c : 0],
                                         dplus[NUP > 1 ? c : 0], uq, acc, vf,
                                         fan);
        }
        __syncthreads();
        if constexpr (fin == fin_local)
          {
            // the result stays at the Gauss points, stored from x-lines
            __builtin_amdgcn_s_setprio(3);
            if (active)
              {
                const int64_t base = c * io.comp_stride + cell * nd;
#pragma unroll
                for (int i = 0; i < n; ++i)
                  io.dst[base + L::template off<0>(p, i)] = io.scale * acc[i];
              }
            __builtin_amdgcn_s_setprio(0);
            continue;
          }
        if constexpr (fin == fin_inverse_mass)
          {
            T ij[n];
            if constexpr (AN)
              {
                // 1/JxW sits at LAST-line points; move it to x-lines
#pragma unroll
                for (int i = 0; i < n; ++i)
                  ij[i] = (T)1 / jxw[i];
                lds_write_line<dim, n, LAST>(g0_l, p, ij);
                __syncthreads();
                lds_read_line<dim, n, 0>(g0_l, p, ij);
              }
            else if constexpr (UNI)
              {
                const T r = (T)(1. / (wl[0] * vol));
#pragma unroll
                for (int i = 0; i < n; ++i)
                  ij[i] = B.winv[i] * r;
              }
            else
#pragma unroll
              for (int i = 0; i < n; ++i)
                ij[i] = ij_row[i];
#pragma unroll
            for (int i = 0; i < n; ++i)
              acc[i] *= ij[i];
          }

        // ---- back to nodal values: x (registers), y, (z) ----
        {
          T t[n];
          if constexpr (fin == fin_inverse_mass)
            eo_apply<n, false, false, false>(B.eo_Sinv, acc, t);
          else
            eo_apply<n, false, true, false>(B.eo_S, acc, t);
          lds_write_line<dim, n, 0>(acc_l, p, t);
          __syncthreads();
          lds_read_line<dim, n, 1>(acc_l, p, t);
          if constexpr (fin == fin_inverse_mass)
            eo_apply<n, false, false, false>(B.eo_Sinv, t, acc);
          else
            eo_apply<n, false, true, false>(B.eo_S, t, acc);
          if constexpr (dim == 3)
            {
              __syncthreads();
              lds_write_line<dim, n, 1>(acc_l, p, acc);
              __syncthreads();
              lds_read_line<dim, n, 2>(acc_l, p, t);
              if constexpr (fin == fin_inverse_mass)
                eo_apply<n, false, false, false>(B.eo_Sinv, t, acc);
              else
                eo_apply<n, false, true, false>(B.eo_S, t, acc);
            }
        }

        // ---- epilogue on LAST-direction lines (coalesced) ----
        __builtin_amdgcn_s_setprio(3);
        if (active)
          {
            const int64_t base = c * io.comp_stride + cell * nd;
            if constexpr (epi == epi_store)
              {
#pragma unroll
                for (int i = 0; i < n; ++i)
                  __builtin_nontemporal_store(
                    io.scale * acc[i],
                    &io.dst[base + L::template off<LAST>(p, i)]);
              }
            else
              {
#pragma unroll
                for (int i = 0; i < n; ++i)
                  {
                    const int64_t k  = base + L::template off<LAST>(p, i);
                    const T       s0 = io.sol_in[k];
                    __builtin_nontemporal_store(fma(io.b_dt, acc[i], s0),
                                                &io.sol_out[k]);
                    if (!io.last)
                      __builtin_nontemporal_store(fma(io.a_dt, acc[i], s0),
                                                  &io.ri_out[k]);
                  }
              }
          }
        if (NC > 1)
          {
            __builtin_amdgcn_s_setprio(0);
            __syncthreads();
          }
      }
  }

  // One batch of CPB cells per workgroup: workgroup b of n_blocks covers
  // the cells of block xcd_block(b, n_blocks). (A grid-stride loop over
  // several batches per workgroup was measured and rejected: the loop lets
  // the compiler keep loop-invariant loads live across iterations, 107 ->
  // 149 VGPRs at 3D degree 4, one occupancy step lost.)
  template <int dim, int n, int NC, typename T, int geo, bool skew,
            bool diff, int fin, int epi, int CPB>
  // the single-stage implicit operator (integrate form, 3D, degree <= 4)
  // at four wavefronts per SIMD: 126-128 VGPRs with at most 4 spilled,
  // instead of 132-146 at three
  __global__ void __launch_bounds__(CPB * (dim == 3 ? n * n : n))
  __attribute__((amdgpu_waves_per_eu(
    dim == 3 && NC == 1 && fin == fin_integrate && !diff && n <= 5 ? 4 : 1)))
  dg_operator_kernel(const OpParams                       P,
                     const DevBasis<T> *__restrict__      Bp,
                     const DevBasis<double> *__restrict__ Bd,
                     const OpBuffers<T>                   io)
  {
    operator_body<dim, n, NC, T, geo, skew, diff, fin, epi, CPB>(
      P, Bp, Bd, io, xcd_block(blockIdx.x, gridDim.x));
  }

  // one member of a batched launch: an independent operator (its own mesh,
  // speed and vectors) of the same variant, owning the workgroups
  // [first_block, first_block + blocks)
  template <typename T>
  struct BatchItem
  {
    OpParams     P;
    OpBuffers<T> io;
    unsigned     first_block;
  };

  // Several operators of one variant in a single launch (the serving loop
  // advances all active solves of one degree by one LSRK stage with it).
  // Items are sorted by first_block; the lookup is uniform per workgroup.
  template <int dim, int n, int NC, typename T, int geo, bool skew,
            bool diff, int fin, int epi, int CPB>
  __global__ void __launch_bounds__(CPB * (dim == 3 ? n * n : n))
  dg_operator_batch_kernel(const BatchItem<T> *__restrict__      items,
                           int                                  n_items,
                           const DevBasis<T> *__restrict__      Bp,
                           const DevBasis<double> *__restrict__ Bd)
  {
    const unsigned b = xcd_block(blockIdx.x, gridDim.x);
    int            r = 0;
    for (int k = 1; k < n_items; ++k)
      if (items[k].first_block <= b)
        r = k;
    const BatchItem<T> &it = items[r];
    operator_body<dim, n, NC, T, geo, skew, diff, fin, epi, CPB>(
      it.P, Bp, Bd, it.io, b - it.first_block);
  }
} // namespace dgops
