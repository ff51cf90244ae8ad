// SIPG diffusion terms of advection_diffusion_solver.cc:531-613 on cells of
// any (multilinear) shape, in integrated form: dst += scale * sign * (the
// diffusion part of the operator tested with every nodal basis function).
// The advection part comes from the operator kernel in its integrate form
// (fin_integrate) and the inverse mass from the line kernel, so the apply
// on such meshes is three launches.
//
// One workgroup per cell. The cell's values, its reference gradients at
// the Gauss points and the integration run as sum-factorized sweeps in LDS;
// the geometry (J, adj J, det J) is evaluated point by point from the 2^dim
// corners. Cell term, with the explicit sign convention (sign = -1):
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
    // stride of the lexicographic node index along direction d
    __device__ static constexpr int
    stride(int d)
    {
      return d == 0 ? 1 : d == 1 ? n : n * n;
    }
  };

  // out[p] (+)= sum_i M[r][i] in[p | r -> i] along direction dir, r the
  // coordinate of p along dir (trans: M[i][r]); optionally a second term
  // with (M2, in2)
  template <int dim, int n, int dir, bool trans, bool two>
  __device__ __forceinline__ void
  sipg_sweep(const double (&M)[max_n][max_n], const double *in,
             const double (&M2)[max_n][max_n], const double *in2, double *out)
  {
    using L           = SipgLay<dim, n>;
    constexpr int st  = L::stride(dir);
    for (int p = threadIdx.x; p < L::nd; p += blockDim.x)
      {
        const int r    = (p / st) % n;
        const int base = p - r * st;
        double    s    = 0;
#pragma unroll
        for (int i = 0; i < n; ++i)
          {
            s = fma(trans ? M[i][r] : M[r][i], in[base + i * st], s);
            if constexpr (two)
              s = fma(trans ? M2[i][r] : M2[r][i], in2[base + i * st], s);
          }
        out[p] = s;
      }
  }

  template <int dim, int n>
  __global__ void __launch_bounds__(128)
  sipg_general_kernel(const OpParams                       P,
                      const DevBasis<double> *__restrict__ Bp,
                      const double *__restrict__           src,
                      double *__restrict__                 dst,
                      double                               scale)
  {
    using L          = SipgLay<dim, n>;
    constexpr int nd = L::nd, nf = L::nf, NV = L::nv;
    __shared__ double su[nd], sn[nd], A[6][nd], sres[nd];
    __shared__ double sX[NV * dim], sXn[NV * dim];
    __shared__ double fw[6][nf]; // tangential partial sums, own | neighbour
    __shared__ double fc[4][nf]; // face coefficients: value, d, t1, t2

    const DevBasis<double> &B    = *Bp;
    const int64_t           cell = blockIdx.x;
    const int               tid  = threadIdx.x;
    const double            nu   = P.viscosity;
    const double            sg   = P.sign;

    for (int p = tid; p < nd; p += blockDim.x)
      {
        su[p]   = src[cell * nd + p];
        sres[p] = 0;
      }
    for (int e = tid; e < NV * dim; e += blockDim.x)
      sX[e] = P.cell_vertices[cell * NV * dim + e];
    __syncthreads();

    // ---- cell term: reference gradients at the Gauss points ----
    int gidx[3]; // A[] slot of the reference gradient along each direction
    if constexpr (dim == 3)
      {
        sipg_sweep<3, n, 0, false, false>(B.S, su, B.S, su, A[0]);
        sipg_sweep<3, n, 0, false, false>(B.G, su, B.G, su, A[1]);
        __syncthreads();
        sipg_sweep<3, n, 1, false, false>(B.S, A[0], B.S, A[0], A[2]);
        sipg_sweep<3, n, 1, false, false>(B.S, A[1], B.S, A[1], A[3]);
        sipg_sweep<3, n, 1, false, false>(B.G, A[0], B.G, A[0], A[4]);
        __syncthreads();
        sipg_sweep<3, n, 2, false, false>(B.G, A[2], B.G, A[2], A[5]);
        sipg_sweep<3, n, 2, false, false>(B.S, A[3], B.S, A[3], A[0]);
        sipg_sweep<3, n, 2, false, false>(B.S, A[4], B.S, A[4], A[1]);
        gidx[0] = 0;
        gidx[1] = 1;
        gidx[2] = 5;
      }
    else
      {
        sipg_sweep<2, n, 0, false, false>(B.S, su, B.S, su, A[0]);
        sipg_sweep<2, n, 0, false, false>(B.G, su, B.G, su, A[1]);
        __syncthreads();
        sipg_sweep<2, n, 1, false, false>(B.S, A[1], B.S, A[1], A[2]);
        sipg_sweep<2, n, 1, false, false>(B.G, A[0], B.G, A[0], A[3]);
        gidx[0] = 2;
        gidx[1] = 3;
        gidx[2] = 3;
      }
    __syncthreads();
    // point-wise: H_d = sign nu w adj(J)_d . J^-T g (in place of g)
    for (int p = tid; p < nd; p += blockDim.x)
      {
        const int q[3] = {p % n, (p / n) % n, dim == 3 ? p / (n * n) : 0};
        double    xi[3] = {B.xq[q[0]], B.xq[q[1]], dim == 3 ? B.xq[q[2]] : 0.};
        double    x[3], J[3][3], adj[3][3];
        multilinear_map<dim>(sX, xi, x, J);
        const double det = adjugate<dim>(J, adj);
        double       w   = B.w[q[0]] * B.w[q[1]];
        if constexpr (dim == 3)
          w *= B.w[q[2]];
        double g[3] = {0, 0, 0}, gu[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < dim; ++d)
          g[d] = A[gidx[d]][p];
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
            A[gidx[d]][p] = f * h;
          }
      }
    __syncthreads();
    // integration with the test-function gradients
    if constexpr (dim == 3)
      {
        sipg_sweep<3, n, 2, true, false>(B.S, A[0], B.S, A[0], A[2]);
        sipg_sweep<3, n, 2, true, false>(B.S, A[1], B.S, A[1], A[3]);
        sipg_sweep<3, n, 2, true, false>(B.G, A[5], B.G, A[5], A[4]);
        __syncthreads();
        sipg_sweep<3, n, 1, true, false>(B.S, A[2], B.S, A[2], A[0]);
        sipg_sweep<3, n, 1, true, true>(B.G, A[3], B.S, A[4], A[1]);
        __syncthreads();
        sipg_sweep<3, n, 0, true, true>(B.G, A[0], B.S, A[1], sres);
      }
    else
      {
        sipg_sweep<2, n, 1, true, false>(B.S, A[2], B.S, A[2], A[0]);
        sipg_sweep<2, n, 1, true, false>(B.G, A[3], B.G, A[3], A[1]);
        __syncthreads();
        sipg_sweep<2, n, 0, true, true>(B.G, A[0], B.S, A[1], sres);
      }
    __syncthreads();

    // ---- faces ----
    for (int f = 0; f < 2 * dim; ++f)
      {
        const int d = f / 2, s = f % 2;
        const int t1 = d == 0 ? 1 : 0, t2 = d == 2 ? 1 : 2;
        const int nb = P.neighbors[cell * 2 * dim + f];
        if (nb < 0) // SIPG runs on periodic meshes; no boundary terms
          continue;
        for (int p = tid; p < nd; p += blockDim.x)
          sn[p] = src[(int64_t)nb * nd + p];
        for (int e = tid; e < NV * dim; e += blockDim.x)
          sXn[e] = P.cell_vertices[(int64_t)nb * NV * dim + e];
        __syncthreads();
        const int sd = L::stride(d), s1 = L::stride(t1),
                  s2 = dim == 3 ? L::stride(t2) : 0;
        // own layer at side s, neighbour layer at side 1 - s: values are the
        // layer's nodes (Gauss-Lobatto), normal derivatives dphi . line;
        // first tangential direction: fw[0..2] own (S v, G v, S dv),
        // fw[3..5] neighbour
        for (int t = tid; t < 2 * nf; t += blockDim.x)
          {
            const int     side = t / nf, pt = t % nf;
            const int     a = pt % n, j = dim == 3 ? pt / n : 0;
            const double *u    = side ? sn : su;
            const int     ls   = side ? 1 - s : s;
            const int     lay  = ls ? n - 1 : 0;
            const double(&dp)[max_n] = ls ? B.dphi1 : B.dphi0;
            double v = 0, vg = 0, dv = 0;
#pragma unroll
            for (int i = 0; i < n; ++i)
              {
                const int base = i * s1 + j * s2;
                const double ui = u[base + lay * sd];
                double       di = 0;
#pragma unroll
                for (int l = 0; l < n; ++l)
                  di = fma(dp[l], u[base + l * sd], di);
                v  = fma(B.S[a][i], ui, v);
                vg = fma(B.G[a][i], ui, vg);
                dv = fma(B.S[a][i], di, dv);
              }
            fw[3 * side][pt]     = v;
            fw[3 * side + 1][pt] = vg;
            fw[3 * side + 2][pt] = dv;
          }
        __syncthreads();
        // second tangential direction, geometry and flux per face point
        for (int pt = tid; pt < nf; pt += blockDim.x)
          {
            const int a = pt % n, b = dim == 3 ? pt / n : 0;
            double    val[2], gref[2][3];
#pragma unroll
            for (int side = 0; side < 2; ++side)
              {
                double v = 0, g1 = 0, g2 = 0, gd = 0;
                if constexpr (dim == 3)
                  {
#pragma unroll
                    for (int jj = 0; jj < n; ++jj)
                      {
                        const int o = jj * n + a;
                        v  = fma(B.S[b][jj], fw[3 * side][o], v);
                        g1 = fma(B.S[b][jj], fw[3 * side + 1][o], g1);
                        g2 = fma(B.G[b][jj], fw[3 * side][o], g2);
                        gd = fma(B.S[b][jj], fw[3 * side + 2][o], gd);
                      }
                  }
                else
                  {
                    v  = fw[3 * side][a];
                    g1 = fw[3 * side + 1][a];
                    gd = fw[3 * side + 2][a];
                  }
                val[side]         = v;
                gref[side][d]     = gd;
                gref[side][t1]    = g1;
                if constexpr (dim == 3)
                  gref[side][t2] = g2;
              }
            double xi[3];
            xi[d]  = s;
            xi[t1] = B.xq[a];
            if constexpr (dim == 3)
              xi[t2] = B.xq[b];
            double x[3], J[3][3], adj[3][3], Jn[3][3], adjn[3][3];
            multilinear_map<dim>(sX, xi, x, J);
            const double det = adjugate<dim>(J, adj);
            xi[d] = 1 - s;
            multilinear_map<dim>(sXn, xi, x, Jn);
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
              nrm[e] = (s ? 1. : -1.) * adj[d][e] / len;
            double JxW = len * B.w[a];
            if constexpr (dim == 3)
              JxW *= B.w[b];
            // J^-1 n on both sides and the normal derivatives
            double jn[3] = {0, 0, 0}, jnn[3] = {0, 0, 0};
#pragma unroll
            for (int k = 0; k < dim; ++k)
              {
#pragma unroll
                for (int e = 0; e < dim; ++e)
                  {
                    jn[k]  = fma(adj[k][e], nrm[e], jn[k]);
                    jnn[k] = fma(adjn[k][e], nrm[e], jnn[k]);
                  }
                jn[k] /= det;
                jnn[k] /= detn;
              }
            double dnm = 0, dnp = 0;
#pragma unroll
            for (int k = 0; k < dim; ++k)
              {
                dnm = fma(jn[k], gref[0][k], dnm);
                dnp = fma(jnn[k], gref[1][k], dnp);
              }
            const double sigma =
              P.penalty * 0.5 * (fabs(jn[d]) + fabs(jnn[d]));
            const double jump    = -0.5 * nu * (val[0] - val[1]);
            const double avg     = -0.5 * nu * (dnm + dnp);
            const double viscous = jump * sigma - avg;
            const double cd      = sg * jump * JxW;
            fc[0][pt] = -sg * viscous * JxW;
            fc[1][pt] = cd * jn[d];
            fc[2][pt] = cd * jn[t1];
            fc[3][pt] = dim == 3 ? cd * jn[t2] : 0.;
          }
        __syncthreads();
        // test functions: along t2 (transposed) into fw, then along t1 and
        // onto the nodes; phi_i(s) picks the layer (Gauss-Lobatto nodes),
        // dphi_s carries the normal-derivative coefficient
        if constexpr (dim == 3)
          {
            for (int t = tid; t < nf; t += blockDim.x)
              {
                const int a = t % n, i2 = t / n;
                double    z0 = 0, z1 = 0, z2 = 0, z3 = 0;
#pragma unroll
                for (int b = 0; b < n; ++b)
                  {
                    const int o = b * n + a;
                    z0 = fma(B.S[b][i2], fc[0][o], z0);
                    z1 = fma(B.S[b][i2], fc[2][o], z1);
                    z2 = fma(B.G[b][i2], fc[3][o], z2);
                    z3 = fma(B.S[b][i2], fc[1][o], z3);
                  }
                fw[0][i2 * n + a] = z0 + z2;
                fw[1][i2 * n + a] = z1;
                fw[2][i2 * n + a] = z3;
              }
            __syncthreads();
          }
        {
          const int      lay = s ? n - 1 : 0;
          const double(&dp)[max_n] = s ? B.dphi1 : B.dphi0;
          for (int p = tid; p < nd; p += blockDim.x)
            {
              const int id = (p / sd) % n, i1 = (p / s1) % n;
              const int i2 = dim == 3 ? (p / s2) % n : 0;
              double    v0 = 0, v1 = 0;
#pragma unroll
              for (int a = 0; a < n; ++a)
                {
                  if constexpr (dim == 3)
                    {
                      const int o = i2 * n + a;
                      v0 = fma(B.S[a][i1], fw[0][o],
                               fma(B.G[a][i1], fw[1][o], v0));
                      v1 = fma(B.S[a][i1], fw[2][o], v1);
                    }
                  else
                    {
                      v0 = fma(B.S[a][i1], fc[0][a],
                               fma(B.G[a][i1], fc[2][a], v0));
                      v1 = fma(B.S[a][i1], fc[1][a], v1);
                    }
                }
              sres[p] += (id == lay ? v0 : 0.) + dp[id] * v1;
            }
        }
        __syncthreads();
      }
    for (int p = tid; p < nd; p += blockDim.x)
      dst[cell * nd + p] += scale * sres[p];
  }
} // namespace dgops
