// Inviscid Burgers right-hand side M^-1(-L(u)) for a dim-component field
// (burgers_solver.cc:259-341): cell flux u (x) u, Rusanov face flux
//   F = 1/2 [(u-.n) u- + (u+.n) u+ + alpha max(|u-.n|, |u+.n|)(u- - u+)],
// cell and face integrals with the over-integration rule QGauss(k(k+1)/2)
// (burgers_solver.cc:234), inverse mass with QGauss(k+1) (:235, :366).
// The flux is nonlinear and the quadrature differs from the nodal points,
// so this kernel works point-wise on LDS tensors (one cell per workgroup,
// one thread per quadrature point) instead of the line-per-thread scheme.
// Faces are element-centric as in op_kernel.hpp: with the cell's own
// outward normal the plus-side flux of the reference is -F.
#pragma once

#include "op_kernel.hpp"

namespace dgops
{
  // over-integration points per direction: k(k+1)/2 <= 28 at degree 7
  constexpr int max_m = 28;

  struct BurgersTables
  {
    int    m;                // over-integration points per direction
    double Sm[max_m][max_n]; // nodal basis at the m Gauss points [q][i]
    double Gm[max_m][max_n]; // its derivative
    double wm[max_m], xm[max_m];
  };

  template <int dim>
  __device__ __forceinline__ int
  ipow_d(int b)
  {
    return dim == 3 ? b * b * b : b * b;
  }

  // out(extent m along dir) = M (m x k, row-major ld) applied along dir to
  // in (extent k along dir); other extents from ext
  __device__ __forceinline__ void
  lds_sweep(const double *M, int ld, bool trans, const double *in, double *out,
            int dir, int m, int k, const int (&ext)[3])
  {
    int eo[3] = {ext[0], ext[1], ext[2]};
    int ei[3] = {ext[0], ext[1], ext[2]};
    eo[dir]   = m;
    ei[dir]   = k;
    const int total = eo[0] * eo[1] * eo[2];
    for (int pt = threadIdx.x; pt < total; pt += blockDim.x)
      {
        int q[3] = {pt % eo[0], (pt / eo[0]) % eo[1], pt / (eo[0] * eo[1])};
        const int qd = q[dir];
        double    s  = 0;
        for (int i = 0; i < k; ++i)
          {
            q[dir]         = i;
            const double a = trans ? M[i * ld + qd] : M[qd * ld + i];
            s += a * in[q[0] + ei[0] * (q[1] + ei[1] * q[2])];
          }
        out[pt] = s;
      }
  }

  template <int dim, int n, int m, typename T, bool cartesian, int epi>
  __global__ void __launch_bounds__(256)
  burgers_kernel(const OpParams P, const DevBasis<double> *__restrict__ Bd,
                 const BurgersTables *__restrict__ BT, const OpBuffers<T> io)
  {
    constexpr int nd  = dim == 3 ? n * n * n : n * n;
    constexpr int md  = dim == 3 ? m * m * m : m * m;
    constexpr int mx  = md > nd ? md : nd;
    constexpr int nf  = dim == 3 ? n * n : n;
    constexpr int mf  = dim == 3 ? m * m : m;
    constexpr int NV  = 1 << dim;
    __shared__ double u_in[dim][nd], res[dim][nd];
    __shared__ double uq[dim][md], vq[dim][md];
    __shared__ double t1[mx], t2[mx], t3[mx];
    __shared__ double fo[dim][mf], fn[dim][mf];
    __shared__ double X[NV * dim];

    const int64_t cell = xcd_block(blockIdx.x, gridDim.x);
    const int     tid  = threadIdx.x;
    const int     nvec = dim * nd;
    const T      *src  = io.src + cell * nvec;
    for (int e = tid; e < NV * dim; e += blockDim.x)
      X[e] = P.cell_vertices[cell * NV * dim + e];
    for (int i = tid; i < nvec; i += blockDim.x)
      {
        u_in[i / nd][i % nd] = (double)src[i];
        res[i / nd][i % nd]  = 0.;
      }
    __syncthreads();

    const double *Sm = &BT->Sm[0][0];
    const double *Gm = &BT->Gm[0][0];
    // values at the m^dim points
    for (int c = 0; c < dim; ++c)
      {
        int ext[3] = {n, n, dim == 3 ? n : 1};
        lds_sweep(Sm, max_n, false, u_in[c], t1, 0, m, n, ext);
        ext[0] = m;
        __syncthreads();
        if (dim == 3)
          {
            lds_sweep(Sm, max_n, false, t1, t2, 1, m, n, ext);
            ext[1] = m;
            __syncthreads();
            lds_sweep(Sm, max_n, false, t2, uq[c], 2, m, n, ext);
          }
        else
          lds_sweep(Sm, max_n, false, t1, uq[c], 1, m, n, ext);
        __syncthreads();
      }
    // v_d = w (adj(J) u)_d, so that J^-1 (u_c u) JxW = u_c v
    for (int pt = tid; pt < md; pt += blockDim.x)
      {
        const int q[3] = {pt % m, (pt / m) % m, dim == 3 ? pt / (m * m) : 0};
        double    w    = BT->wm[q[0]] * BT->wm[q[1]];
        if (dim == 3)
          w *= BT->wm[q[2]];
        double adj[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        if (cartesian)
          {
            double vol = P.h[0] * P.h[1] * (dim == 3 ? P.h[2] : 1.);
            for (int d = 0; d < dim; ++d)
              adj[d][d] = vol / P.h[d];
          }
        else
          {
            const double xi[3] = {BT->xm[q[0]], BT->xm[q[1]],
                                  dim == 3 ? BT->xm[q[2]] : 0.};
            double       x[3], J[3][3];
            multilinear_map<dim>(X, xi, x, J);
            adjugate<dim>(J, adj);
          }
        for (int d = 0; d < dim; ++d)
          {
            double s = 0;
            for (int e = 0; e < dim; ++e)
              s += adj[d][e] * uq[e][pt];
            vq[d][pt] = w * s;
          }
      }
    __syncthreads();
    // cell term: res_c += sum_d (G along d, S elsewhere)^T (u_c v_d)
    for (int c = 0; c < dim; ++c)
      for (int d = 0; d < dim; ++d)
        {
          for (int pt = tid; pt < md; pt += blockDim.x)
            t1[pt] = uq[c][pt] * vq[d][pt];
          __syncthreads();
          int ext[3] = {m, m, dim == 3 ? m : 1};
          lds_sweep(d == 0 ? Gm : Sm, max_n, true, t1, t2, 0, n, m, ext);
          ext[0] = n;
          __syncthreads();
          lds_sweep(d == 1 ? Gm : Sm, max_n, true, t2, t3, 1, n, m, ext);
          ext[1] = n;
          __syncthreads();
          double *r = t3;
          if (dim == 3)
            {
              lds_sweep(d == 2 ? Gm : Sm, max_n, true, t3, t2, 2, n, m, ext);
              __syncthreads();
              r = t2;
            }
          // -L: the cell integral enters with + (burgers_solver.cc:280-284)
          for (int i = tid; i < nd; i += blockDim.x)
            res[c][i] += r[i];
          __syncthreads();
        }
    // faces
    for (int f = 0; f < 2 * dim; ++f)
      {
        const int d = f / 2, s = f % 2;
        const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
        const int nb = P.neighbors[cell * 2 * dim + f];
        // own and neighbour face layers -> face quadrature points
        for (int c = 0; c < dim; ++c)
          for (int side = 0; side < 2; ++side)
            {
              for (int j = tid; j < nf; j += blockDim.x)
                {
                  int ijk[3] = {0, 0, 0};
                  ijk[ta]    = j % n;
                  if (dim == 3)
                    ijk[tb] = j / n;
                  ijk[d]            = s ? n - 1 : 0;
                  const int idx_own = ijk[0] + n * (ijk[1] + n * ijk[2]);
                  ijk[d]            = s ? 0 : n - 1;
                  const int idx_nb  = ijk[0] + n * (ijk[1] + n * ijk[2]);
                  t1[j] = side == 0 ?
                            u_in[c][idx_own] :
                            (double)io.src[(int64_t)nb * nvec + c * nd + idx_nb];
                }
              __syncthreads();
              int ext[3] = {n, dim == 3 ? n : 1, 1};
              lds_sweep(Sm, max_n, false, t1, t2, 0, m, n, ext);
              ext[0] = m;
              __syncthreads();
              double *dstf = side == 0 ? fo[c] : fn[c];
              if (dim == 3)
                lds_sweep(Sm, max_n, false, t2, dstf, 1, m, n, ext);
              else
                for (int j = tid; j < mf; j += blockDim.x)
                  dstf[j] = t2[j];
              __syncthreads();
            }
        // flux at the face points (outward normal of this cell)
        for (int fq = tid; fq < mf; fq += blockDim.x)
          {
            const int a = fq % m, bq = dim == 3 ? fq / m : 0;
            double    w = BT->wm[a] * (dim == 3 ? BT->wm[bq] : 1.);
            double    nrm[3] = {0, 0, 0};
            if (cartesian)
              {
                double area = 1;
                for (int e = 0; e < dim; ++e)
                  if (e != d)
                    area *= P.h[e];
                nrm[d] = (s ? 1. : -1.) * area * w;
              }
            else
              {
                double xi[3] = {0, 0, 0};
                xi[d]  = s;
                xi[ta] = BT->xm[a];
                if (dim == 3)
                  xi[tb] = BT->xm[bq];
                double x[3], J[3][3], adj[3][3];
                multilinear_map<dim>(X, xi, x, J);
                adjugate<dim>(J, adj);
                for (int e = 0; e < dim; ++e)
                  nrm[e] = (s ? 1. : -1.) * adj[d][e] * w;
              }
            // nrm = n dA; unit-normal quantities via |n dA|
            double len = 0;
            for (int e = 0; e < dim; ++e)
              len += nrm[e] * nrm[e];
            len = sqrt(len);
            double unm = 0, unp = 0;
            for (int e = 0; e < dim; ++e)
              {
                unm += fo[e][fq] * nrm[e] / len;
                unp += fn[e][fq] * nrm[e] / len;
              }
            const double lam = P.alpha * fmax(fabs(unm), fabs(unp));
            double       flux[3];
            for (int c = 0; c < dim; ++c)
              flux[c] = 0.5 * (unm * fo[c][fq] + unp * fn[c][fq] +
                               lam * (fo[c][fq] - fn[c][fq])) *
                        len;
            for (int c = 0; c < dim; ++c)
              fo[c][fq] = -flux[c];
          }
        __syncthreads();
        // integrate onto the own face layer
        for (int c = 0; c < dim; ++c)
          {
            int ext[3] = {m, dim == 3 ? m : 1, 1};
            lds_sweep(Sm, max_n, true, fo[c], t1, 0, n, m, ext);
            ext[0] = n;
            __syncthreads();
            double *lay = t1;
            if (dim == 3)
              {
                lds_sweep(Sm, max_n, true, t1, t2, 1, n, m, ext);
                __syncthreads();
                lay = t2;
              }
            for (int j = tid; j < nf; j += blockDim.x)
              {
                int ijk[3] = {0, 0, 0};
                ijk[ta]    = j % n;
                if (dim == 3)
                  ijk[tb] = j / n;
                ijk[d] = s ? n - 1 : 0;
                res[c][ijk[0] + n * (ijk[1] + n * ijk[2])] += lay[j];
              }
            __syncthreads();
          }
      }
    // inverse mass with the (k+1)-point rule: S^-1 diag(1/JxW) S^-T
    for (int c = 0; c < dim; ++c)
      {
        const double *Si = &Bd->Sinv[0][0];
        int           ext[3] = {n, n, dim == 3 ? n : 1};
        lds_sweep(Si, max_n, true, res[c], t1, 0, n, n, ext);
        __syncthreads();
        lds_sweep(Si, max_n, true, t1, t2, 1, n, n, ext);
        __syncthreads();
        double *a = t2, *b = t1;
        if (dim == 3)
          {
            lds_sweep(Si, max_n, true, t2, t1, 2, n, n, ext);
            __syncthreads();
            a = t1;
            b = t2;
          }
        for (int pt = tid; pt < nd; pt += blockDim.x)
          {
            const int q[3] = {pt % n, (pt / n) % n, dim == 3 ? pt / (n * n) : 0};
            double    w = Bd->w[q[0]] * Bd->w[q[1]] * (dim == 3 ? Bd->w[q[2]] : 1.);
            double    det;
            if (cartesian)
              det = P.h[0] * P.h[1] * (dim == 3 ? P.h[2] : 1.);
            else
              {
                const double xi[3] = {Bd->xq[q[0]], Bd->xq[q[1]],
                                      dim == 3 ? Bd->xq[q[2]] : 0.};
                double       x[3], J[3][3], adj[3][3];
                multilinear_map<dim>(X, xi, x, J);
                det = adjugate<dim>(J, adj);
              }
            a[pt] /= det * w;
          }
        __syncthreads();
        lds_sweep(Si, max_n, false, a, b, 0, n, n, ext);
        __syncthreads();
        lds_sweep(Si, max_n, false, b, a, 1, n, n, ext);
        __syncthreads();
        double *r = a;
        if (dim == 3)
          {
            lds_sweep(Si, max_n, false, a, b, 2, n, n, ext);
            __syncthreads();
            r = b;
          }
        for (int i = tid; i < nd; i += blockDim.x)
          {
            const int64_t k = cell * nvec + c * nd + i;
            if constexpr (epi == epi_store)
              io.dst[k] = io.scale * (T)r[i];
            else
              {
                const T s0    = io.sol_in[k];
                io.sol_out[k] = fma(io.b_dt, (T)r[i], s0);
                if (!io.last)
                  io.ri_out[k] = fma(io.a_dt, (T)r[i], s0);
              }
          }
        __syncthreads();
      }
  }
} // namespace dgops
