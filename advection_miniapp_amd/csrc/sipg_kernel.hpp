// SIPG diffusion terms of advection_diffusion_solver.cc:531-613 on cells of
// any (multilinear) shape, in integrated form: dst += scale * sign * (the
// diffusion part of the operator tested with every nodal basis function).
// The advection part comes from the operator kernel in its integrate form
// (fin_integrate) and the inverse mass from the line kernel, so the apply
// on such meshes is three launches.
//
// CPB cells per workgroup of 256 threads, each cell with its own block of
// LDS; every phase strides its tasks (a point, face point or node of one of
// the workgroup's cells) over the threads. The cell's values, its reference
// gradients at the Gauss points and the integration run as sum-factorized
// sweeps in LDS; the cell term's geometry (J, adj J, det J) is evaluated
// point by point from the 2^dim corners, the faces' from a table built once
// per mesh (sipg_face_table_kernel). Cell term, with the explicit sign convention
// (sign = -1):
//   sign * int grad v . nu grad u = sum_q grad_ref v . (sign nu w_q adj(J)
//   J^-T grad_ref u).
// Faces are element-centric: each cell evaluates its own and the
// neighbour's trace and reference gradient at the face's Gauss points and
// adds its own side of
//   jump = -nu/2 (u- - u+), avg = -nu/2 (d_n u- + d_n u+),
//   value term -sign (jump sigma - avg) JxW, normal-derivative term
//   sign jump JxW (n . J^-T grad_ref v),
// with its own outward normal n; sigma = penalty (|(J-^-1 n)_d| +
// |(J+^-1 n)_d|) / 2 per face point (d the face's reference direction).
// The reference evaluates sigma once per face batch from the exterior cell
// at the first face point (:576-581); on affine meshes the two are equal.
#pragma once

#include "dev_basis.hpp"
#include "dg_common.hpp"
#include "line_geometry.hpp"

namespace dgops
{
  template <int dim, int n>
  struct SipgLay
  {
    static constexpr int nd = dim == 3 ? n * n * n : n * n;
    static constexpr int nf = dim == 3 ? n * n : n;
    static constexpr int nv = 1 << dim;
    // face table values per face point: J^-1 n of the cell below the face
    // (dim), of the cell above it (dim), JxW
    static constexpr int nt = 2 * dim + 1;
    // stride of the lexicographic node index along direction d
    __device__ static constexpr int
    stride(int d)
    {
      return d == 0 ? 1 : d == 1 ? n : n * n;
    }
    // per-cell LDS block (doubles): own values, the neighbours' values
    // across the two faces of a direction, six point buffers of the cell
    // term (in the face passes: the two faces' geometry [face s][nt][nf] as
    // seen from this cell), the result, the corners, the face partial sums
    // [face s][own | neighbour][3 nf] and the face coefficients
    // [face s][4][nf]
    static constexpr int oU   = 0;
    static constexpr int oN   = oU + nd;
    static constexpr int oA   = oN + 2 * nd;
    static constexpr int oG   = oA;
    static constexpr int nA   = 6 * nd > 2 * nt * nf ? 6 * nd : 2 * nt * nf;
    static constexpr int oR   = oA + nA;
    static constexpr int oX   = oR + nd;
    static constexpr int oFW  = oX + nv * dim;
    static constexpr int oFC  = oFW + 12 * nf;
    static constexpr int size = oFC + 8 * nf;
    // threads per workgroup and cells per workgroup: one point task per
    // thread where the cells allow it, at most 48 KB of LDS
    static constexpr int NT  = 256;
    static constexpr int cpb_t = nd >= NT ? 1 : NT / nd;
    static constexpr int cpb_l = 48 * 1024 / (8 * size);
    static constexpr int CPB   = cpb_l < 1 ? 1 : cpb_t < cpb_l ? cpb_t : cpb_l;
  };

  // out[p] (+)= sum_i M[r][i] in[p | r -> i] along direction dir, r the
  // coordinate of p along dir (trans: M[i][r]); optionally a second term
  // with (M2, in2). in, in2, out: offsets in the cells' LDS blocks
  template <int dim, int n, int dir, bool trans, bool two>
  __device__ __forceinline__ void
  sipg_sweep(double *lds, const double *M, int in, const double *M2, int in2,
             int out)
  {
    using L          = SipgLay<dim, n>;
    constexpr int st = L::stride(dir);
    for (int t = threadIdx.x; t < L::CPB * L::nd; t += L::NT)
      {
        const int     lc   = t / L::nd, p = t - lc * L::nd;
        double       *c    = lds + lc * L::size;
        const int     r    = (p / st) % n;
        const int     base = p - r * st;
        double        s    = 0;
#pragma unroll
        for (int i = 0; i < n; ++i)
          {
            s = fma(trans ? M[i * n + r] : M[r * n + i], c[in + base + i * st],
                    s);
            if constexpr (two)
              s = fma(trans ? M2[i * n + r] : M2[r * n + i],
                      c[in2 + base + i * st], s);
          }
        c[out + p] = s;
      }
  }

  template <int d>
  struct FaceDir
  {
    static constexpr int value = d;
  };

  // Geometry of the face above each cell in each direction d (its face
  // xi_d = 1, the neighbour's xi_d = 0), at the face's Gauss points:
  // T[((cell dim + d) nt + v) nf + pt] = J^-1 n of the cell (v < dim), of
  // the neighbour (dim <= v < 2 dim), JxW (v = 2 dim), with the cell's unit
  // outward normal n. Seen from the neighbour the two J^-1 n exchange and
  // change sign, which gives every cell the face below it. Only interior
  // vertices are perturbed, so faces across periodic boundaries match.
  template <int dim, int n, int d>
  __device__ __forceinline__ void
  sipg_face_entry(const OpParams &P, const DevBasis<double> &B, int64_t cell,
                  int pt, double *__restrict__ T)
  {
    using L          = SipgLay<dim, n>;
    constexpr int nf = L::nf, NV = L::nv;
    constexpr int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
    const int     a = pt % n, b = dim == 3 ? pt / n : 0;
    double        xi[3] = {0, 0, 0};
    xi[t1] = B.xq[a];
    if constexpr (dim == 3)
      xi[t2] = B.xq[b];
    double x[3], J[3][3], adj[3][3], Jn[3][3], adjn[3][3];
    xi[d] = 1;
    multilinear_map<dim>(P.cell_vertices + cell * NV * dim, xi, x, J);
    const double det = adjugate<dim>(J, adj);
    const int    nb  = P.neighbors[cell * 2 * dim + 2 * d + 1];
    xi[d] = 0;
    multilinear_map<dim>(P.cell_vertices + (int64_t)(nb >= 0 ? nb : cell) *
                                             NV * dim,
                         xi, x, Jn);
    const double detn = adjugate<dim>(Jn, adjn);
    // area vector of the face (row d of adj J, along +xi_d)
    double len = 0;
#pragma unroll
    for (int e = 0; e < dim; ++e)
      len = fma(adj[d][e], adj[d][e], len);
    len = sqrt(len);
    double nrm[3] = {0, 0, 0};
#pragma unroll
    for (int e = 0; e < dim; ++e)
      nrm[e] = adj[d][e] / len;
    double JxW = len * B.w[a];
    if constexpr (dim == 3)
      JxW *= B.w[b];
    double *t = T + ((cell * dim + d) * L::nt) * nf + pt;
#pragma unroll
    for (int k = 0; k < dim; ++k)
      {
        double jn = 0, jnn = 0;
#pragma unroll
        for (int e = 0; e < dim; ++e)
          {
            jn  = fma(adj[k][e], nrm[e], jn);
            jnn = fma(adjn[k][e], nrm[e], jnn);
          }
        t[k * nf]         = jn / det;
        t[(dim + k) * nf] = jnn / detn;
      }
    t[2 * dim * nf] = JxW;
  }

  template <int dim, int n>
  __global__ void __launch_bounds__(256)
  sipg_face_table_kernel(const OpParams                       P,
                         const DevBasis<double> *__restrict__ Bp,
                         double *__restrict__                 T)
  {
    constexpr int nf = SipgLay<dim, n>::nf;
    const int64_t e  = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= P.n_cells * dim * nf)
      return;
    const int     pt   = (int)(e % nf);
    const int64_t cd   = e / nf;
    const int     d    = (int)(cd % dim);
    const int64_t cell = cd / dim;
    if (d == 0)
      sipg_face_entry<dim, n, 0>(P, *Bp, cell, pt, T);
    else if (d == 1)
      sipg_face_entry<dim, n, 1>(P, *Bp, cell, pt, T);
    else if constexpr (dim == 3)
      sipg_face_entry<dim, n, 2>(P, *Bp, cell, pt, T);
  }

  template <int dim, int n>
  __global__ void __launch_bounds__(256)
  sipg_general_kernel(const OpParams                       P,
                      const DevBasis<double> *__restrict__ Bp,
                      const double *__restrict__           faces,
                      const double *__restrict__           src,
                      double *__restrict__                 dst,
                      double                               scale)
  {
    using L           = SipgLay<dim, n>;
    constexpr int nd  = L::nd, nf = L::nf, NV = L::nv, NT = L::NT;
    constexpr int CPB = L::CPB, SZ = L::size;
    __shared__ double lds[CPB * SZ];
    __shared__ int    nbr[CPB][2 * dim];
    // the 1D tables (row stride n): S and G at the Gauss points, and the
    // nodal-basis derivatives at the two faces
    __shared__ double tS[n * n], tG[n * n], tD[2][n];

    const DevBasis<double> &B     = *Bp;
    const int64_t           cell0 = (int64_t)blockIdx.x * CPB;
    const int               tid   = threadIdx.x;
    const double            nu    = P.viscosity;
    const double            sg    = P.sign;
    auto cell_of = [&](int lc) { return cell0 + lc; };
    auto active  = [&](int lc) { return cell0 + lc < P.n_cells; };
    auto A       = [](int k) { return L::oA + k * nd; };

    for (int t = tid; t < CPB * nd; t += NT)
      {
        const int lc = t / nd, p = t - lc * nd;
        double   *c  = lds + lc * SZ;
        c[L::oU + p] = active(lc) ? src[cell_of(lc) * nd + p] : 0.;
        c[L::oR + p] = 0;
      }
    for (int t = tid; t < CPB * NV * dim; t += NT)
      {
        const int lc = t / (NV * dim), e = t - lc * NV * dim;
        lds[lc * SZ + L::oX + e] =
          active(lc) ? P.cell_vertices[cell_of(lc) * NV * dim + e] : 0.;
      }
    for (int t = tid; t < CPB * 2 * dim; t += NT)
      {
        const int lc = t / (2 * dim), f = t - lc * 2 * dim;
        nbr[lc][f]   = active(lc) ? P.neighbors[cell_of(lc) * 2 * dim + f] : -1;
      }
    for (int e = tid; e < n * n; e += NT)
      {
        tS[e] = B.S[e / n][e % n];
        tG[e] = B.G[e / n][e % n];
      }
    for (int e = tid; e < 2 * n; e += NT)
      tD[e / n][e % n] = e < n ? B.dphi0[e] : B.dphi1[e - n];
    __syncthreads();

    // ---- cell term: reference gradients at the Gauss points ----
    int gidx[3]; // A() slot of the reference gradient along each direction
    if constexpr (dim == 3)
      {
        sipg_sweep<3, n, 0, false, false>(lds, tS, L::oU, tS, L::oU, A(0));
        sipg_sweep<3, n, 0, false, false>(lds, tG, L::oU, tG, L::oU, A(1));
        __syncthreads();
        sipg_sweep<3, n, 1, false, false>(lds, tS, A(0), tS, A(0), A(2));
        sipg_sweep<3, n, 1, false, false>(lds, tS, A(1), tS, A(1), A(3));
        sipg_sweep<3, n, 1, false, false>(lds, tG, A(0), tG, A(0), A(4));
        __syncthreads();
        sipg_sweep<3, n, 2, false, false>(lds, tG, A(2), tG, A(2), A(5));
        sipg_sweep<3, n, 2, false, false>(lds, tS, A(3), tS, A(3), A(0));
        sipg_sweep<3, n, 2, false, false>(lds, tS, A(4), tS, A(4), A(1));
        gidx[0] = 0;
        gidx[1] = 1;
        gidx[2] = 5;
      }
    else
      {
        sipg_sweep<2, n, 0, false, false>(lds, tS, L::oU, tS, L::oU, A(0));
        sipg_sweep<2, n, 0, false, false>(lds, tG, L::oU, tG, L::oU, A(1));
        __syncthreads();
        sipg_sweep<2, n, 1, false, false>(lds, tS, A(1), tS, A(1), A(2));
        sipg_sweep<2, n, 1, false, false>(lds, tG, A(0), tG, A(0), A(3));
        gidx[0] = 2;
        gidx[1] = 3;
        gidx[2] = 3;
      }
    __syncthreads();
    // point-wise: H_d = sign nu w adj(J)_d . J^-T g (in place of g)
    for (int t = tid; t < CPB * nd; t += NT)
      {
        const int lc = t / nd, p = t - lc * nd;
        double   *c  = lds + lc * SZ;
        const int q[3] = {p % n, (p / n) % n, dim == 3 ? p / (n * n) : 0};
        double    xi[3] = {B.xq[q[0]], B.xq[q[1]], dim == 3 ? B.xq[q[2]] : 0.};
        double    x[3], J[3][3], adj[3][3];
        multilinear_map<dim>(c + L::oX, xi, x, J);
        const double det = adjugate<dim>(J, adj);
        double       w   = B.w[q[0]] * B.w[q[1]];
        if constexpr (dim == 3)
          w *= B.w[q[2]];
        double g[3] = {0, 0, 0}, gu[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < dim; ++d)
          g[d] = c[A(gidx[d]) + p];
#pragma unroll
        for (int e = 0; e < dim; ++e)
#pragma unroll
          for (int d = 0; d < dim; ++d)
            gu[e] = fma(adj[d][e], g[d], gu[e]);
        const double f = sg * nu * w / det;
#pragma unroll
        for (int d = 0; d < dim; ++d)
          {
            double h = 0;
#pragma unroll
            for (int e = 0; e < dim; ++e)
              h = fma(adj[d][e], gu[e], h);
            c[A(gidx[d]) + p] = f * h;
          }
      }
    __syncthreads();
    // integration with the test-function gradients
    if constexpr (dim == 3)
      {
        sipg_sweep<3, n, 2, true, false>(lds, tS, A(0), tS, A(0), A(2));
        sipg_sweep<3, n, 2, true, false>(lds, tS, A(1), tS, A(1), A(3));
        sipg_sweep<3, n, 2, true, false>(lds, tG, A(5), tG, A(5), A(4));
        __syncthreads();
        sipg_sweep<3, n, 1, true, false>(lds, tS, A(2), tS, A(2), A(0));
        sipg_sweep<3, n, 1, true, true>(lds, tG, A(3), tS, A(4), A(1));
        __syncthreads();
        sipg_sweep<3, n, 0, true, true>(lds, tG, A(0), tS, A(1), L::oR);
      }
    else
      {
        sipg_sweep<2, n, 1, true, false>(lds, tS, A(2), tS, A(2), A(0));
        sipg_sweep<2, n, 1, true, false>(lds, tG, A(3), tG, A(3), A(1));
        __syncthreads();
        sipg_sweep<2, n, 0, true, true>(lds, tG, A(0), tS, A(1), L::oR);
      }
    __syncthreads();

    // ---- faces: both faces of a direction per pass (direction d a
    // compile-time constant, so every index below is static) ----
    auto face_pair = [&](auto dc) {
      constexpr int d = decltype(dc)::value;
      constexpr int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
      constexpr int sd = L::stride(d), s1 = L::stride(t1),
                    s2 = dim == 3 ? L::stride(t2) : 0;
      // the neighbours across the two faces (periodic meshes: both exist)
      for (int t = tid; t < CPB * 2 * nd; t += NT)
        {
          const int lc = t / (2 * nd), e = t - lc * 2 * nd;
          const int s = e / nd, q = e - s * nd;
          const int nb = nbr[lc][2 * d + s];
          lds[lc * SZ + L::oN + e] = nb >= 0 ? src[(int64_t)nb * nd + q] : 0.;
        }
      // the geometry of the two faces from the table: face 1 is the entry
      // of this cell, face 0 the entry of the neighbour below, with the two
      // J^-1 n exchanged and negated (SIPG meshes are periodic: the
      // neighbour exists)
      constexpr int nt = L::nt;
      for (int t = tid; t < CPB * 2 * nf; t += NT)
        {
          const int lc = t / (2 * nf), e = t - lc * 2 * nf;
          const int s = e / nf, pt = e - s * nf;
          if (!active(lc))
            continue;
          const int64_t owner = s ? cell_of(lc) : (int64_t)nbr[lc][2 * d];
          const double *f     = faces + (owner * dim + d) * nt * nf + pt;
          double       *g     = lds + lc * SZ + L::oG + s * nt * nf + pt;
#pragma unroll
          for (int k = 0; k < dim; ++k)
            {
              const double own = f[k * nf], other = f[(dim + k) * nf];
              g[k * nf]         = s ? own : -other;
              g[(dim + k) * nf] = s ? other : -own;
            }
          g[2 * dim * nf] = f[2 * dim * nf];
        }
      __syncthreads();
      // layer values and normal derivatives, first tangential direction:
      // FW[s][side][0..2] = (S v, G v, S dv) at (a, j); side 0 own (layer
      // s), side 1 the neighbour across face s (its layer 1 - s)
      for (int t = tid; t < CPB * 4 * nf; t += NT)
        {
          const int     lc = t / (4 * nf), e = t - lc * 4 * nf;
          double       *c  = lds + lc * SZ;
          const int     s = e / (2 * nf), side = (e / nf) % 2, pt = e % nf;
          const int     a = pt % n, j = dim == 3 ? pt / n : 0;
          const double *u  = side ? c + L::oN + s * nd : c + L::oU;
          const int     ls = side ? 1 - s : s;
          const int     lay = ls ? n - 1 : 0;
          const double *dp = tD[ls];
          double v = 0, vg = 0, dv = 0;
#pragma unroll
          for (int i = 0; i < n; ++i)
            {
              const int    base = i * s1 + j * s2;
              const double ui   = u[base + lay * sd];
              double       di   = 0;
#pragma unroll
              for (int l = 0; l < n; ++l)
                di = fma(dp[l], u[base + l * sd], di);
              v  = fma(tS[a * n + i], ui, v);
              vg = fma(tG[a * n + i], ui, vg);
              dv = fma(tS[a * n + i], di, dv);
            }
          double *o = c + L::oFW + (2 * s + side) * 3 * nf;
          o[pt]          = v;
          o[nf + pt]     = vg;
          o[2 * nf + pt] = dv;
        }
      __syncthreads();
      // second tangential direction and the flux per face point
      for (int t = tid; t < CPB * 2 * nf; t += NT)
        {
          const int lc = t / (2 * nf), e = t - lc * 2 * nf;
          double   *c  = lds + lc * SZ;
          const int s = e / nf, pt = e % nf;
          const int a = pt % n, b = dim == 3 ? pt / n : 0;
          double    val[2], gref[2][3];
#pragma unroll
          for (int side = 0; side < 2; ++side)
            {
              const double *o = c + L::oFW + (2 * s + side) * 3 * nf;
              double        v = 0, g1 = 0, g2 = 0, gd = 0;
              if constexpr (dim == 3)
                {
#pragma unroll
                  for (int jj = 0; jj < n; ++jj)
                    {
                      const int q = jj * n + a;
                      v  = fma(tS[b * n + jj], o[q], v);
                      g1 = fma(tS[b * n + jj], o[nf + q], g1);
                      g2 = fma(tG[b * n + jj], o[q], g2);
                      gd = fma(tS[b * n + jj], o[2 * nf + q], gd);
                    }
                }
              else
                {
                  v  = o[a];
                  g1 = o[nf + a];
                  gd = o[2 * nf + a];
                }
              val[side]      = v;
              gref[side][d]  = gd;
              gref[side][t1] = g1;
              if constexpr (dim == 3)
                gref[side][t2] = g2;
            }
          // J^-1 n with this cell's outward normal on both sides, JxW
          const double *g = c + L::oG + s * nt * nf + pt;
          double        jn[3] = {0, 0, 0}, jnn[3] = {0, 0, 0};
#pragma unroll
          for (int k = 0; k < dim; ++k)
            {
              jn[k]  = g[k * nf];
              jnn[k] = g[(dim + k) * nf];
            }
          const double JxW = g[2 * dim * nf];
          double dnm = 0, dnp = 0;
#pragma unroll
          for (int k = 0; k < dim; ++k)
            {
              dnm = fma(jn[k], gref[0][k], dnm);
              dnp = fma(jnn[k], gref[1][k], dnp);
            }
          const double sigma   = P.penalty * 0.5 * (fabs(jn[d]) + fabs(jnn[d]));
          const double jump    = -0.5 * nu * (val[0] - val[1]);
          const double avg     = -0.5 * nu * (dnm + dnp);
          const double viscous = jump * sigma - avg;
          const double cd      = sg * jump * JxW;
          double      *fc      = c + L::oFC + s * 4 * nf;
          fc[pt]          = -sg * viscous * JxW;
          fc[nf + pt]     = cd * jn[d];
          fc[2 * nf + pt] = cd * jn[t1];
          fc[3 * nf + pt] = dim == 3 ? cd * jn[t2] : 0.;
        }
      __syncthreads();
      // test functions: along t2 (transposed) into FW[s][0], then along t1
      // and onto the nodes; phi_i(s) picks the layer (Gauss-Lobatto nodes),
      // dphi_s carries the normal-derivative coefficient
      if constexpr (dim == 3)
        {
          for (int t = tid; t < CPB * 2 * nf; t += NT)
            {
              const int     lc = t / (2 * nf), e = t - lc * 2 * nf;
              double       *c  = lds + lc * SZ;
              const int     s = e / nf, pt = e % nf;
              const int     a = pt % n, i2 = pt / n;
              const double *fc = c + L::oFC + s * 4 * nf;
              double        z0 = 0, z1 = 0, z2 = 0, z3 = 0;
#pragma unroll
              for (int b = 0; b < n; ++b)
                {
                  const int o = b * n + a;
                  z0 = fma(tS[b * n + i2], fc[o], z0);
                  z1 = fma(tS[b * n + i2], fc[2 * nf + o], z1);
                  z2 = fma(tG[b * n + i2], fc[3 * nf + o], z2);
                  z3 = fma(tS[b * n + i2], fc[nf + o], z3);
                }
              double *o = c + L::oFW + 2 * s * 3 * nf;
              o[pt]          = z0 + z2;
              o[nf + pt]     = z1;
              o[2 * nf + pt] = z3;
            }
          __syncthreads();
        }
      for (int t = tid; t < CPB * nd; t += NT)
        {
          const int lc = t / nd, p = t - lc * nd;
          double   *c  = lds + lc * SZ;
          const int id = (p / sd) % n, i1 = (p / s1) % n;
          const int i2 = dim == 3 ? (p / s2) % n : 0;
          double    add = 0;
#pragma unroll
          for (int s = 0; s < 2; ++s)
            {
              const double *dp = tD[s];
              const double *o  = c + L::oFW + 2 * s * 3 * nf;
              const double *fc = c + L::oFC + s * 4 * nf;
              double        v0 = 0, v1 = 0;
#pragma unroll
              for (int a = 0; a < n; ++a)
                {
                  if constexpr (dim == 3)
                    {
                      const int q = i2 * n + a;
                      v0 = fma(tS[a * n + i1], o[q],
                               fma(tG[a * n + i1], o[nf + q], v0));
                      v1 = fma(tS[a * n + i1], o[2 * nf + q], v1);
                    }
                  else
                    {
                      v0 = fma(tS[a * n + i1], fc[a],
                               fma(tG[a * n + i1], fc[2 * nf + a], v0));
                      v1 = fma(tS[a * n + i1], fc[nf + a], v1);
                    }
                }
              add += (id == (s ? n - 1 : 0) ? v0 : 0.) + dp[id] * v1;
            }
          c[L::oR + p] += add;
        }
      __syncthreads();
    };
    face_pair(FaceDir<0>{});
    face_pair(FaceDir<1>{});
    if constexpr (dim == 3)
      face_pair(FaceDir<2>{});
    for (int t = tid; t < CPB * nd; t += NT)
      {
        const int lc = t / nd, p = t - lc * nd;
        if (active(lc))
          dst[cell_of(lc) * nd + p] += scale * lds[lc * SZ + L::oR + p];
      }
  }
} // namespace dgops
