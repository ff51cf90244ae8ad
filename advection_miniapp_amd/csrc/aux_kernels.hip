// Auxiliary kernels: L2 projection of analytic data, cell integrals
// (mass, energy, H1 seminorm, L2 error), vector operations and the
// pointwise updates of the time integrators. These run once per output or
// are plain streaming kernels; the hot operator lives in op_kernel.hpp.
#include <map>
#include <mutex>
#include <string>

#include "aux.hpp"
#include "op_kernel.hpp"
#include "registry.hpp"

namespace dgops
{
  namespace
  {
    constexpr int aux_threads = 256;
    constexpr int max_pts     = 729; // (max_n + 1)^3

    struct QuadTables
    {
      int    n2;             // points of the second rule
      double S2[9][max_n];   // nodal basis at the n2-point Gauss rule
      double w2[9], x2[9];
      // the n-point Gauss rule, also where the operator integrates with
      // Gauss-Lobatto points (use_gl_quad)
      double w1[max_n], x1[max_n];
    };

    // tensor-product sweep on LDS data: out has extent m along `dir`,
    // in has extent k along `dir`, other extents given by ext[]
    __device__ void
    sweep(const double *M, int ld, const double *in, double *out, int dir,
          int m, int k, const int (&ext)[3])
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
              q[dir] = i;
              s += M[qd * ld + i] * in[q[0] + ei[0] * (q[1] + ei[1] * q[2])];
            }
          out[pt] = s;
        }
    }

    template <typename T>
    __device__ T
    load_as(const void *p, int64_t i)
    {
      return static_cast<const T *>(p)[i];
    }

    template <int dim>
    __global__ void
    project_kernel(const OpParams P, const DevBasis<double> *__restrict__ B,
                   int n, int nc, double t, void *dst, int fp32)
    {
      __shared__ double a[max_pts], b[max_pts];
      __shared__ double X[24];
      constexpr int     NV = 1 << dim;
      const int64_t     cell = blockIdx.x;
      for (int e = threadIdx.x; e < NV * dim; e += blockDim.x)
        X[e] = P.cell_vertices[cell * NV * dim + e];
      __syncthreads();
      const int nd     = dim == 3 ? n * n * n : n * n;
      int       ext[3] = {n, n, dim == 3 ? n : 1};
      for (int c = 0; c < nc; ++c)
        {
          for (int pt = threadIdx.x; pt < nd; pt += blockDim.x)
            {
              const int q[3] = {pt % n, (pt / n) % n, dim == 3 ? pt / (n * n) : 0};
              double xi[3] = {B->xq[q[0]], B->xq[q[1]], dim == 3 ? B->xq[q[2]] : 0.};
              double x[3], J[3][3];
              multilinear_map<dim>(X, xi, x, J);
              a[pt] = analytic_solution<dim>(P.sol_kind, P.sol_param, t,
                                             P.vel_const, x, c);
            }
          __syncthreads();
          // S^-1 in every direction (transform_from_q_points_to_basis)
          sweep(&B->Sinv[0][0], max_n, a, b, 0, n, n, ext);
          __syncthreads();
          sweep(&B->Sinv[0][0], max_n, b, a, 1, n, n, ext);
          __syncthreads();
          double *res = a;
          if (dim == 3)
            {
              sweep(&B->Sinv[0][0], max_n, a, b, 2, n, n, ext);
              __syncthreads();
              res = b;
            }
          for (int pt = threadIdx.x; pt < nd; pt += blockDim.x)
            {
              const int64_t k = (cell * nc + c) * nd + pt;
              if (fp32)
                static_cast<float *>(dst)[k] = (float)res[pt];
              else
                static_cast<double *>(dst)[k] = res[pt];
            }
          __syncthreads();
        }
    }

    __device__ void
    block_sum(double *red, double v, double *out)
    {
      red[threadIdx.x] = v;
      __syncthreads();
      for (int s = blockDim.x / 2; s > 0; s >>= 1)
        {
          if ((int)threadIdx.x < s)
            red[threadIdx.x] += red[threadIdx.x + s];
          __syncthreads();
        }
      if (threadIdx.x == 0)
        *out = red[0];
      __syncthreads();
    }

    // per cell: mass, energy, H1 seminorm (advection_solver.cc:787-818)
    template <int dim>
    __global__ void
    mass_energy_kernel(const OpParams P, const DevBasis<double> *__restrict__ B,
                       int n, const void *vec, int fp32, double *partials)
    {
      __shared__ double a[max_pts], b[max_pts];
      __shared__ double X[24];
      __shared__ double red[aux_threads];
      constexpr int     NV   = 1 << dim;
      const int64_t     cell = blockIdx.x;
      const int         nd   = dim == 3 ? n * n * n : n * n;
      int               ext[3] = {n, n, dim == 3 ? n : 1};
      for (int e = threadIdx.x; e < NV * dim; e += blockDim.x)
        X[e] = P.cell_vertices[cell * NV * dim + e];
      for (int pt = threadIdx.x; pt < nd; pt += blockDim.x)
        a[pt] = fp32 ? (double)load_as<float>(vec, cell * nd + pt) :
                       load_as<double>(vec, cell * nd + pt);
      __syncthreads();
      sweep(&B->S[0][0], max_n, a, b, 0, n, n, ext);
      __syncthreads();
      sweep(&B->S[0][0], max_n, b, a, 1, n, n, ext);
      __syncthreads();
      double *uq = a;
      if (dim == 3)
        {
          sweep(&B->S[0][0], max_n, a, b, 2, n, n, ext);
          __syncthreads();
          uq = b;
        }
      double m = 0, en = 0, h1 = 0;
      for (int pt = threadIdx.x; pt < nd; pt += blockDim.x)
        {
          const int q[3] = {pt % n, (pt / n) % n, dim == 3 ? pt / (n * n) : 0};
          double    xi[3] = {B->xq[q[0]], B->xq[q[1]], dim == 3 ? B->xq[q[2]] : 0.};
          double    x[3], J[3][3], adj[3][3];
          multilinear_map<dim>(X, xi, x, J);
          const double det = adjugate<dim>(J, adj);
          double       w   = B->w[q[0]] * B->w[q[1]] * (dim == 3 ? B->w[q[2]] : 1.);
          const double JxW = det * w;
          double       gref[3] = {0, 0, 0};
          for (int d = 0; d < dim; ++d)
            {
              int qq[3] = {q[0], q[1], q[2]};
              for (int p = 0; p < n; ++p)
                {
                  qq[d] = p;
                  gref[d] += B->D[q[d]][p] * uq[qq[0] + n * (qq[1] + n * qq[2])];
                }
            }
          double g2 = 0;
          for (int e = 0; e < dim; ++e)
            {
              double ge = 0;
              for (int d = 0; d < dim; ++d)
                ge += adj[d][e] * gref[d];
              ge /= det;
              g2 += ge * ge;
            }
          const double u = uq[pt];
          m += u * JxW;
          en += u * u * JxW;
          h1 += g2 * JxW;
        }
      block_sum(red, m, &partials[cell * 3 + 0]);
      block_sum(red, en, &partials[cell * 3 + 1]);
      block_sum(red, h1, &partials[cell * 3 + 2]);
    }

    // per cell: squared L2 error with the (n+1)-point rule and squared
    // norm of the analytic solution with the n-point Gauss rule
    // (advection_solver.cc:1187-1208), whatever the operator's quadrature
    template <int dim>
    __global__ void
    l2_error_kernel(const OpParams P, const DevBasis<double> *__restrict__ B,
                    const QuadTables Q, int n, const void *vec, int fp32,
                    double t, double *partials)
    {
      __shared__ double a[max_pts], b[max_pts];
      __shared__ double X[24];
      __shared__ double red[aux_threads];
      constexpr int     NV   = 1 << dim;
      const int64_t     cell = blockIdx.x;
      const int         nd   = dim == 3 ? n * n * n : n * n;
      const int         n2   = Q.n2;
      for (int e = threadIdx.x; e < NV * dim; e += blockDim.x)
        X[e] = P.cell_vertices[cell * NV * dim + e];
      for (int pt = threadIdx.x; pt < nd; pt += blockDim.x)
        a[pt] = fp32 ? (double)load_as<float>(vec, cell * nd + pt) :
                       load_as<double>(vec, cell * nd + pt);
      __syncthreads();
      int ext[3] = {n, n, dim == 3 ? n : 1};
      sweep(&Q.S2[0][0], max_n, a, b, 0, n2, n, ext);
      ext[0] = n2;
      __syncthreads();
      sweep(&Q.S2[0][0], max_n, b, a, 1, n2, n, ext);
      ext[1] = n2;
      __syncthreads();
      double *uq = a;
      if (dim == 3)
        {
          sweep(&Q.S2[0][0], max_n, a, b, 2, n2, n, ext);
          __syncthreads();
          uq = b;
        }
      const int nd2 = dim == 3 ? n2 * n2 * n2 : n2 * n2;
      double    err = 0, nrm = 0;
      for (int pt = threadIdx.x; pt < nd2; pt += blockDim.x)
        {
          const int q[3] = {pt % n2, (pt / n2) % n2, dim == 3 ? pt / (n2 * n2) : 0};
          double    xi[3] = {Q.x2[q[0]], Q.x2[q[1]], dim == 3 ? Q.x2[q[2]] : 0.};
          double    x[3], J[3][3], adj[3][3];
          multilinear_map<dim>(X, xi, x, J);
          const double det = adjugate<dim>(J, adj);
          const double w = Q.w2[q[0]] * Q.w2[q[1]] * (dim == 3 ? Q.w2[q[2]] : 1.);
          const double g = analytic_solution<dim>(P.sol_kind, P.sol_param, t,
                                                  P.vel_const, x);
          err += (uq[pt] - g) * (uq[pt] - g) * det * w;
        }
      for (int pt = threadIdx.x; pt < nd; pt += blockDim.x)
        {
          const int q[3] = {pt % n, (pt / n) % n, dim == 3 ? pt / (n * n) : 0};
          double    xi[3] = {Q.x1[q[0]], Q.x1[q[1]], dim == 3 ? Q.x1[q[2]] : 0.};
          double    x[3], J[3][3], adj[3][3];
          multilinear_map<dim>(X, xi, x, J);
          const double det = adjugate<dim>(J, adj);
          const double w = Q.w1[q[0]] * Q.w1[q[1]] * (dim == 3 ? Q.w1[q[2]] : 1.);
          const double g = analytic_solution<dim>(P.sol_kind, P.sol_param, t,
                                                  P.vel_const, x);
          nrm += g * g * det * w;
        }
      block_sum(red, err, &partials[cell * 2 + 0]);
      block_sum(red, nrm, &partials[cell * 2 + 1]);
    }

    template <int dim>
    __global__ void
    fdm_cell_data_kernel(const OpParams P,
                         const DevBasis<double> *__restrict__ B, int n,
                         double *velocity, double *invdet)
    {
      __shared__ double X[24];
      __shared__ double red[aux_threads];
      constexpr int     NV   = 1 << dim;
      const int64_t     cell = blockIdx.x;
      const int         nd   = dim == 3 ? n * n * n : n * n;
      for (int e = threadIdx.x; e < NV * dim; e += blockDim.x)
        X[e] = P.cell_vertices[cell * NV * dim + e];
      __syncthreads();
      double va[3] = {0, 0, 0}, vol = 0;
      for (int pt = threadIdx.x; pt < nd; pt += blockDim.x)
        {
          const int q[3] = {pt % n, (pt / n) % n, dim == 3 ? pt / (n * n) : 0};
          double    xi[3] = {B->xq[q[0]], B->xq[q[1]], dim == 3 ? B->xq[q[2]] : 0.};
          double    x[3], J[3][3], adj[3][3], a[3] = {0, 0, 0};
          multilinear_map<dim>(X, xi, x, J);
          const double det = adjugate<dim>(J, adj);
          velocity_field<dim>(P.vel_kind, P.vel_const, x, a);
          const double w =
            B->w[q[0]] * B->w[q[1]] * (dim == 3 ? B->w[q[2]] : 1.);
          // (J^-1 a) JxW = adj(J) a w
          for (int d = 0; d < dim; ++d)
            {
              double s = 0;
              for (int e = 0; e < dim; ++e)
                s += adj[d][e] * a[e];
              va[d] += s * w;
            }
          vol += det * w;
          if (pt == 0)
            invdet[cell] = 1. / det;
        }
      double tot[3], v;
      for (int d = 0; d < dim; ++d)
        block_sum(red, va[d], &tot[d]);
      block_sum(red, vol, &v);
      if (threadIdx.x == 0)
        for (int d = 0; d < dim; ++d)
          velocity[cell * dim + d] = tot[d] / v;
    }

    // sum of `k` interleaved columns of partials[count][k]
    __global__ void
    reduce_columns(const double *partials, int64_t count, int k, double *out)
    {
      __shared__ double red[1024];
      for (int c = 0; c < k; ++c)
        {
          double s = 0;
          for (int64_t i = threadIdx.x; i < count; i += blockDim.x)
            s += partials[i * k + c];
          red[threadIdx.x] = s;
          __syncthreads();
          for (int st = blockDim.x / 2; st > 0; st >>= 1)
            {
              if ((int)threadIdx.x < st)
                red[threadIdx.x] += red[threadIdx.x + st];
              __syncthreads();
            }
          if (threadIdx.x == 0)
            out[c] = red[0];
          __syncthreads();
        }
    }

    // ---- vector kernels ----------------------------------------------------

    constexpr int vec_threads = 256;
    constexpr int max_multi   = 16;

    struct Coeffs
    {
      double v[max_multi];
    };

    template <typename T>
    __global__ void
    axpby_kernel(int64_t n, T a, const T *__restrict__ x, T b, T *y)
    {
      for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
           i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (b == (T)0) ? a * x[i] : fma(a, x[i], b * y[i]);
    }

    // y = (y + sum_j c_j x_j) * scale
    template <typename T, int M>
    __global__ void
    multi_axpy_kernel(int64_t n, const Coeffs c, const T *__restrict__ x,
                      int64_t xs, T *y, T scale)
    {
      for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
           i += (int64_t)gridDim.x * blockDim.x)
        {
          T s = y[i];
#pragma unroll
          for (int j = 0; j < M; ++j)
            s = fma((T)c.v[j], x[j * xs + i], s);
          y[i] = s * scale;
        }
    }

    // column sums of partials[count][k] for count <= 1024: one wavefront
    // per column, no barrier (the block partials of the vector kernels)
    __global__ void
    reduce_columns_small(const double *partials, int count, int k,
                         double *out)
    {
      const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
      for (int c = wave; c < k; c += blockDim.x / 64)
        {
          double s = 0;
          for (int i = lane; i < count; i += 64)
            s += partials[i * k + c];
          for (int off = 32; off > 0; off >>= 1)
            s += __shfl_down(s, off, 64);
          if (lane == 0)
            out[c] = s;
        }
    }

    template <typename T, int M>
    __global__ void
    multi_dot_kernel(int64_t n, const T *__restrict__ x, int64_t xs,
                     const T *__restrict__ y, double *partials)
    {
      __shared__ double red[M][vec_threads / 64];
      double            s[M];
#pragma unroll
      for (int j = 0; j < M; ++j)
        s[j] = 0;
      for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
           i += (int64_t)gridDim.x * blockDim.x)
        {
          const double yi = (double)y[i];
#pragma unroll
          for (int j = 0; j < M; ++j)
            s[j] = fma((double)x[j * xs + i], yi, s[j]);
        }
      const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
      for (int j = 0; j < M; ++j)
        {
          double v = s[j];
          for (int off = 32; off > 0; off >>= 1)
            v += __shfl_down(v, off, 64);
          if (lane == 0)
            red[j][wave] = v;
        }
      __syncthreads();
      if (threadIdx.x < M)
        {
          double v = 0;
          for (int w = 0; w < vec_threads / 64; ++w)
            v += red[threadIdx.x][w];
          partials[blockIdx.x * M + threadIdx.x] = v;
        }
    }

    // y += sum_j c_j x_j, then <x_j, y> (j < M) and <y, y> of the updated
    // y in the same pass: the subtraction of one Gram-Schmidt sweep fused
    // with the projections of the next
    constexpr int max_fused = 32;
    struct FusedCoeffs
    {
      double v[max_fused];
    };

    template <typename T, int M>
    __global__ void __launch_bounds__(vec_threads)
    multi_axpy_dot_kernel(int64_t n, const FusedCoeffs c,
                          const T *__restrict__ x, int64_t xs, T *y,
                          double *partials)
    {
      __shared__ double red[M + 1][vec_threads / 64];
      double            s[M + 1];
#pragma unroll
      for (int j = 0; j <= M; ++j)
        s[j] = 0;
      for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
           i += (int64_t)gridDim.x * blockDim.x)
        {
          T xv[M];
          T yi = y[i];
#pragma unroll
          for (int j = 0; j < M; ++j)
            {
              xv[j] = x[j * xs + i];
              yi    = fma((T)c.v[j], xv[j], yi);
            }
          y[i] = yi;
          const double yd = (double)yi;
#pragma unroll
          for (int j = 0; j < M; ++j)
            s[j] = fma((double)xv[j], yd, s[j]);
          s[M] = fma(yd, yd, s[M]);
        }
      const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
      for (int j = 0; j <= M; ++j)
        {
          double v = s[j];
          for (int off = 32; off > 0; off >>= 1)
            v += __shfl_down(v, off, 64);
          if (lane == 0)
            red[j][wave] = v;
        }
      __syncthreads();
      if (threadIdx.x <= M)
        {
          double v = 0;
          for (int w = 0; w < vec_threads / 64; ++w)
            v += red[threadIdx.x][w];
          partials[blockIdx.x * (M + 1) + threadIdx.x] = v;
        }
    }

    // ---- Arnoldi steps with their scalars in device memory ------------------
    // (the fixed-length inner GMRES of the mixed-precision solve runs with
    // no host round trip)

    constexpr int max_arnoldi = 16;

    // x *= 1 / sqrt(*s2) (x = 0 when *s2 = 0: then x was zero already)
    template <typename T>
    __global__ void
    scale_inv_sqrt_kernel(int64_t n, const double *__restrict__ s2, T *x)
    {
      const double v = *s2;
      const T      f = v > 0 ? (T)(1. / sqrt(v)) : (T)0;
      for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
           i += (int64_t)gridDim.x * blockDim.x)
        x[i] *= f;
    }

    // norm of w - sum_k h_k V_k from h[0..m-1] = <V_k, w> and h[m] = <w, w>
    // (one classical Gram-Schmidt sweep against orthonormal V_k)
    __device__ __forceinline__ double
    arnoldi_norm(const double *__restrict__ h, int m)
    {
      double hh = 0;
      for (int k = 0; k < m; ++k)
        hh = fma(h[k], h[k], hh);
      const double r2 = h[m] - hh;
      return r2 > 0 ? sqrt(r2) : 0.;
    }

    // w = (w - sum_k h_k V_k) / arnoldi_norm(h, M) (unscaled on breakdown)
    template <typename T, int M>
    __global__ void
    arnoldi_update_kernel(int64_t n, const double *__restrict__ h,
                          const T *__restrict__ V, int64_t vs, T *w)
    {
      T c[M];
#pragma unroll
      for (int k = 0; k < M; ++k)
        c[k] = (T)h[k];
      const double hn = arnoldi_norm(h, M);
      const T      f  = hn > 0 ? (T)(1. / hn) : (T)1;
      for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
           i += (int64_t)gridDim.x * blockDim.x)
        {
          T s = w[i];
#pragma unroll
          for (int k = 0; k < M; ++k)
            s = fma(-c[k], V[k * vs + i], s);
          w[i] = s * f;
        }
    }

    // min |beta e1 - H y| for the (m+1) x m Hessenberg matrix whose column
    // j sits at h + j hs as (H_0j .. H_jj, <w_j, w_j>), the subdiagonal
    // H_j+1,j = arnoldi_norm of that column; Givens QR and back
    // substitution by one thread
    __global__ void
    arnoldi_lsq_kernel(int m, const double *__restrict__ h, int64_t hs,
                       const double *__restrict__ beta2, double *y)
    {
      __shared__ double H[(max_arnoldi + 1) * max_arnoldi];
      __shared__ double g[max_arnoldi + 1], cs[max_arnoldi],
        sn[max_arnoldi];
      if (threadIdx.x != 0)
        return;
      for (int j = 0; j < m; ++j)
        {
          const double *col = h + j * hs;
          for (int k = 0; k <= m; ++k)
            H[k * max_arnoldi + j] =
              k <= j ? col[k] : k == j + 1 ? arnoldi_norm(col, j + 1) : 0.;
        }
      const double b2 = *beta2;
      g[0]            = b2 > 0 ? sqrt(b2) : 0.;
      for (int k = 1; k <= m; ++k)
        g[k] = 0;
      for (int j = 0; j < m; ++j)
        {
          for (int k = 0; k < j; ++k)
            {
              const double a = H[k * max_arnoldi + j],
                           b = H[(k + 1) * max_arnoldi + j];
              H[k * max_arnoldi + j]       = cs[k] * a + sn[k] * b;
              H[(k + 1) * max_arnoldi + j] = -sn[k] * a + cs[k] * b;
            }
          const double a = H[j * max_arnoldi + j],
                       b = H[(j + 1) * max_arnoldi + j];
          const double r = sqrt(a * a + b * b);
          cs[j]          = r == 0 ? 1. : a / r;
          sn[j]          = r == 0 ? 0. : b / r;
          H[j * max_arnoldi + j] = r;
          g[j + 1]               = -sn[j] * g[j];
          g[j]                   = cs[j] * g[j];
        }
      for (int k = m - 1; k >= 0; --k)
        {
          double s = g[k];
          for (int l = k + 1; l < m; ++l)
            s -= H[k * max_arnoldi + l] * y[l];
          const double d = H[k * max_arnoldi + k];
          y[k]           = d != 0 ? s / d : 0.;
        }
    }

    // x = sum_k y_k Z_k, accumulated in FP64, stored as TO
    template <typename TI, typename TO, int M>
    __global__ void
    combine_kernel(int64_t n, const double *__restrict__ y,
                   const TI *__restrict__ Z, int64_t zs, TO *x)
    {
      double c[M];
#pragma unroll
      for (int k = 0; k < M; ++k)
        c[k] = y[k];
      for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
           i += (int64_t)gridDim.x * blockDim.x)
        {
          double s = 0;
#pragma unroll
          for (int k = 0; k < M; ++k)
            s = fma(c[k], (double)Z[k * zs + i], s);
          x[i] = (TO)s;
        }
    }

    __global__ void
    gram4_kernel(int64_t n, const double *__restrict__ v0,
                 const double *__restrict__ v1, const double *__restrict__ v2,
                 const double *__restrict__ v3, const double *__restrict__ r,
                 double *partials)
    {
      __shared__ double red[14][vec_threads / 64];
      double            s[14];
      for (int j = 0; j < 14; ++j)
        s[j] = 0;
      for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
           i += (int64_t)gridDim.x * blockDim.x)
        {
          const double a[4] = {v0[i], v1[i], v2[i], v3[i]};
          const double b    = r[i];
          int          c    = 0;
          for (int l = 0; l < 4; ++l)
            for (int j = 0; j <= l; ++j, ++c)
              s[c] = fma(a[l], a[j], s[c]);
          for (int l = 0; l < 4; ++l, ++c)
            s[c] = fma(a[l], b, s[c]);
        }
      const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
      for (int j = 0; j < 14; ++j)
        {
          double v = s[j];
          for (int off = 32; off > 0; off >>= 1)
            v += __shfl_down(v, off, 64);
          if (lane == 0)
            red[j][wave] = v;
        }
      __syncthreads();
      if (threadIdx.x < 14)
        {
          double v = 0;
          for (int w = 0; w < vec_threads / 64; ++w)
            v += red[threadIdx.x][w];
          partials[blockIdx.x * 14 + threadIdx.x] = v;
        }
    }

    template <typename TS, typename TD>
    __global__ void
    convert_kernel(int64_t n, const TS *__restrict__ x, TD *__restrict__ y)
    {
      for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
           i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (TD)x[i];
    }

    template <typename T>
    __global__ void
    lsrk_update_kernel(int64_t n, T b, T a, const T *k, const T *sol_in,
                       T *sol_out, T *ri_out)
    {
      for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
           i += (int64_t)gridDim.x * blockDim.x)
        {
          const T ki = k[i];
          const T si = sol_in[i];
          sol_out[i] = fma(b, ki, si);
          if (ri_out)
            ri_out[i] = fma(a, ki, si);
        }
    }

    unsigned
    grid_for(int64_t n)
    {
      const int64_t b = (n + vec_threads - 1) / vec_threads;
      return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 4096));
    }

    template <typename T>
    void
    multi_dot_t(Operator &op, int64_t n, int m, const T *x, int64_t xs,
                const T *y, double *result)
    {
      const unsigned nb = std::min<unsigned>(grid_for(n), 1024);
      op.d_partials.alloc(sizeof(double) * nb * max_multi);
      // one sum per vector x_j: room for all m of them
      op.d_result.alloc(sizeof(double) * std::max(64, m));
      double *part = op.d_partials.as<double>();
      int     done = 0;
      std::vector<double> host(m);
      while (done < m)
        {
          const int chunk = std::min(max_multi, m - done);
          const T  *xc    = x + done * xs;
#define DG_MD(M)                                                              \
  case M:                                                                     \
    hipLaunchKernelGGL((multi_dot_kernel<T, M>), dim3(nb), dim3(vec_threads), \
                       0, op.stream, n, xc, xs, y, part);                     \
    break;
          switch (chunk)
            {
              DG_MD(1) DG_MD(2) DG_MD(3) DG_MD(4) DG_MD(5) DG_MD(6) DG_MD(7)
              DG_MD(8) DG_MD(9) DG_MD(10) DG_MD(11) DG_MD(12) DG_MD(13)
              DG_MD(14) DG_MD(15) DG_MD(16)
            }
#undef DG_MD
          hip_check(hipGetLastError(), "multi_dot");
          hipLaunchKernelGGL(reduce_columns_small, dim3(1), dim3(1024), 0,
                             op.stream, part, (int)nb, chunk,
                             op.d_result.as<double>() + done);
          hip_check(hipGetLastError(), "reduce");
          done += chunk;
        }
      hip_check(hipMemcpyAsync(host.data(), op.d_result.ptr,
                               sizeof(double) * m, hipMemcpyDeviceToHost,
                               op.stream),
                "copy result");
      hip_check(hipStreamSynchronize(op.stream), "sync");
      for (int j = 0; j < m; ++j)
        result[j] = host[j];
    }

    // multi_dot_t with the sums left in device memory at d_out[0..m-1]
    template <typename T>
    void
    multi_dot_dev_t(Operator &op, int64_t n, int m, const T *x, int64_t xs,
                    const T *y, double *d_out)
    {
      const unsigned nb = std::min<unsigned>(grid_for(n), 1024);
      op.d_partials.alloc(sizeof(double) * nb * max_multi);
      double *part = op.d_partials.as<double>();
      int     done = 0;
      while (done < m)
        {
          const int chunk = std::min(max_multi, m - done);
          const T  *xc    = x + done * xs;
#define DG_MD(M)                                                              \
  case M:                                                                     \
    hipLaunchKernelGGL((multi_dot_kernel<T, M>), dim3(nb), dim3(vec_threads), \
                       0, op.stream, n, xc, xs, y, part);                     \
    break;
          switch (chunk)
            {
              DG_MD(1) DG_MD(2) DG_MD(3) DG_MD(4) DG_MD(5) DG_MD(6) DG_MD(7)
              DG_MD(8) DG_MD(9) DG_MD(10) DG_MD(11) DG_MD(12) DG_MD(13)
              DG_MD(14) DG_MD(15) DG_MD(16)
            }
#undef DG_MD
          hip_check(hipGetLastError(), "multi_dot");
          hipLaunchKernelGGL(reduce_columns_small, dim3(1), dim3(1024), 0,
                             op.stream, part, (int)nb, chunk, d_out + done);
          hip_check(hipGetLastError(), "reduce");
          done += chunk;
        }
    }

    template <typename T>
    void
    arnoldi_update_t(Operator &op, int64_t n, int m, const double *d_h,
                     const T *V, int64_t vs, T *w)
    {
#define DG_AU(M)                                                            \
  case M:                                                                   \
    hipLaunchKernelGGL((arnoldi_update_kernel<T, M>), dim3(grid_for(n)),    \
                       dim3(vec_threads), 0, op.stream, n, d_h, V, vs, w);  \
    break;
      switch (m)
        {
          DG_AU(1) DG_AU(2) DG_AU(3) DG_AU(4) DG_AU(5) DG_AU(6) DG_AU(7)
          DG_AU(8) DG_AU(9) DG_AU(10) DG_AU(11) DG_AU(12) DG_AU(13)
          DG_AU(14) DG_AU(15) DG_AU(16)
        }
#undef DG_AU
      hip_check(hipGetLastError(), "arnoldi update");
    }

    template <typename TI>
    void
    combine_t(Operator &op, int64_t n, int m, const double *d_y, const TI *Z,
              int64_t zs, double *x)
    {
#define DG_CB(M)                                                              \
  case M:                                                                     \
    hipLaunchKernelGGL((combine_kernel<TI, double, M>), dim3(grid_for(n)),    \
                       dim3(vec_threads), 0, op.stream, n, d_y, Z, zs, x);    \
    break;
      switch (m)
        {
          DG_CB(1) DG_CB(2) DG_CB(3) DG_CB(4) DG_CB(5) DG_CB(6) DG_CB(7)
          DG_CB(8) DG_CB(9) DG_CB(10) DG_CB(11) DG_CB(12) DG_CB(13)
          DG_CB(14) DG_CB(15) DG_CB(16)
        }
#undef DG_CB
      hip_check(hipGetLastError(), "combine");
    }

    template <typename T>
    void
    multi_axpy_dot_t(Operator &op, int64_t n, int m, const double *alpha,
                     const T *x, int64_t xs, T *y, double *result)
    {
      const unsigned nb = std::min<unsigned>(grid_for(n), 1024);
      op.d_partials.alloc(sizeof(double) * nb * (max_fused + 1));
      op.d_result.alloc(sizeof(double) * 64);
      FusedCoeffs c = {};
      for (int j = 0; j < m; ++j)
        c.v[j] = alpha[j];
      double *part = op.d_partials.as<double>();
#define DG_MAD(M)                                                           \
  case M:                                                                   \
    hipLaunchKernelGGL((multi_axpy_dot_kernel<T, M>), dim3(nb),             \
                       dim3(vec_threads), 0, op.stream, n, c, x, xs, y,     \
                       part);                                               \
    break;
      switch (m)
        {
          DG_MAD(1) DG_MAD(2) DG_MAD(3) DG_MAD(4) DG_MAD(5) DG_MAD(6)
          DG_MAD(7) DG_MAD(8) DG_MAD(9) DG_MAD(10) DG_MAD(11) DG_MAD(12)
          DG_MAD(13) DG_MAD(14) DG_MAD(15) DG_MAD(16) DG_MAD(17) DG_MAD(18)
          DG_MAD(19) DG_MAD(20) DG_MAD(21) DG_MAD(22) DG_MAD(23) DG_MAD(24)
          DG_MAD(25) DG_MAD(26) DG_MAD(27) DG_MAD(28) DG_MAD(29) DG_MAD(30)
          DG_MAD(31) DG_MAD(32)
        }
#undef DG_MAD
      hip_check(hipGetLastError(), "multi_axpy_dot");
      hipLaunchKernelGGL(reduce_columns_small, dim3(1), dim3(1024), 0,
                         op.stream, part, (int)nb, m + 1,
                         op.d_result.as<double>());
      hip_check(hipGetLastError(), "reduce");
      std::vector<double> host(m + 1);
      hip_check(hipMemcpyAsync(host.data(), op.d_result.ptr,
                               sizeof(double) * (m + 1),
                               hipMemcpyDeviceToHost, op.stream),
                "copy result");
      hip_check(hipStreamSynchronize(op.stream), "sync");
      for (int j = 0; j <= m; ++j)
        result[j] = host[j];
    }

    // y = (y + sum_j alpha_j x_j) * scale, in chunks of max_multi vectors
    // (the scale goes with the last chunk)
    template <typename T>
    void
    multi_axpy_t(Operator &op, int64_t n, int m, const double *alpha,
                 const T *x, int64_t xs, T *y, double scale = 1.)
    {
      int done = 0;
      while (done < m)
        {
          const int chunk = std::min(max_multi, m - done);
          Coeffs    c     = {};
          for (int j = 0; j < chunk; ++j)
            c.v[j] = alpha[done + j];
          const T *xc = x + done * xs;
          const T  sc = done + chunk == m ? (T)scale : (T)1;
#define DG_MA(M)                                                             \
  case M:                                                                    \
    hipLaunchKernelGGL((multi_axpy_kernel<T, M>), dim3(grid_for(n)),         \
                       dim3(vec_threads), 0, op.stream, n, c, xc, xs, y, sc);\
    break;
          switch (chunk)
            {
              DG_MA(1) DG_MA(2) DG_MA(3) DG_MA(4) DG_MA(5) DG_MA(6) DG_MA(7)
              DG_MA(8) DG_MA(9) DG_MA(10) DG_MA(11) DG_MA(12) DG_MA(13)
              DG_MA(14) DG_MA(15) DG_MA(16)
            }
#undef DG_MA
          hip_check(hipGetLastError(), "multi_axpy");
          done += chunk;
        }
    }

    QuadTables
    make_quad_tables(int n)
    {
      QuadTables Q = {};
      Q.n2         = n + 1;
      std::vector<double> gp, gw;
      gauss_points(n, gp, gw);
      for (int q = 0; q < n; ++q)
        {
          Q.w1[q] = gw[q];
          Q.x1[q] = gp[q];
        }
      gauss_points(n + 1, gp, gw);
      const auto xl = gauss_lobatto_points_ext(n);
      for (int q = 0; q < n + 1; ++q)
        {
          Q.w2[q] = gw[q];
          Q.x2[q] = gp[q];
          for (int i = 0; i < n; ++i)
            {
              real_ext v, d;
              lagrange(xl, i, (real_ext)gp[q], v, d);
              Q.S2[q][i] = (double)v;
            }
        }
      return Q;
    }

    // ---- geometry tables of the operator kernels ---------------------------

    template <int dim, int n>
    __device__ void
    face_line_point(int d, int p, int (&q)[3])
    {
      if (d == 0)
        Lay<dim, n>::template idx<0>(p, 0, q);
      else if (d == 1)
        Lay<dim, n>::template idx<1>(p, 0, q);
      else
        Lay<dim, n>::template idx<dim - 1>(p, 0, q);
    }

    // one block per cell: w adj(J) a and 1/JxW at the Gauss points
    // (lexicographic), and w (adj(J) a)_d on the lower face of every
    // direction, indexed by the line p of direction d that ends on it
    template <int dim, int n>
    __global__ void
    geometry_table_kernel(const OpParams P,
                          const DevBasis<double> *__restrict__ B,
                          double *tab_cell, double *tab_face)
    {
      constexpr int NV = 1 << dim;
      constexpr int nl = Lay<dim, n>::nl, nd = Lay<dim, n>::nd;
      __shared__ double X[NV * dim];
      const int64_t     cell = blockIdx.x;
      for (int e = threadIdx.x; e < NV * dim; e += blockDim.x)
        X[e] = P.cell_vertices[cell * NV * dim + e];
      __syncthreads();
      double *tc = tab_cell + cell * (dim + 1) * nd;
      for (int pt = threadIdx.x; pt < nd; pt += blockDim.x)
        {
          const int q[3] = {pt % n, (pt / n) % n,
                            dim == 3 ? pt / (n * n) : 0};
          double    cv[3], jxw;
          cell_point_data<dim>(P, X, B->xq, B->w, q, cv, jxw);
          // row d in the line order of direction d (row dim, read along x,
          // in that of x): point i of line p at i * nl + p
          for (int d = 0; d < dim; ++d)
            {
              const int i = q[d];
              const int p = d == 0 ? q[1] + n * q[2]
                          : d == 1 ? q[0] + n * q[2]
                                   : q[0] + n * q[1];
              tc[d * nd + i * nl + p] = cv[d];
            }
          tc[dim * nd + q[0] * nl + q[1] + n * q[2]] = 1. / jxw;
        }
      for (int t = threadIdx.x; t < dim * nl; t += blockDim.x)
        {
          const int d = t / nl, p = t % nl;
          int       q[3];
          face_line_point<dim, n>(d, p, q);
          double an, x[3];
          face_point_data<dim>(P, X, B->xq, B->w, d, 0, q, an, x);
          tab_face[(cell * dim + d) * nl + p] = -an;
        }
    }

    // one block per boundary face: outward a . n dA and the physical point
    template <int dim, int n>
    __global__ void
    boundary_table_kernel(const OpParams P,
                          const DevBasis<double> *__restrict__ B,
                          const int *bcell, const int *bface, double *tab)
    {
      constexpr int NV = 1 << dim;
      constexpr int nl = Lay<dim, n>::nl;
      __shared__ double X[NV * dim];
      const int64_t     b    = blockIdx.x;
      const int64_t     cell = bcell[b];
      const int         f    = bface[b];
      for (int e = threadIdx.x; e < NV * dim; e += blockDim.x)
        X[e] = P.cell_vertices[cell * NV * dim + e];
      __syncthreads();
      for (int p = threadIdx.x; p < nl; p += blockDim.x)
        {
          int q[3];
          face_line_point<dim, n>(f / 2, p, q);
          double an, x[3] = {0, 0, 0};
          face_point_data<dim>(P, X, B->xq, B->w, f / 2, f % 2, q, an,
                                      x);
          double *r = tab + (b * nl + p) * (1 + dim);
          r[0]      = an;
          for (int e = 0; e < dim; ++e)
            r[1 + e] = x[e];
        }
    }

    template <int dim>
    __global__ void
    boundary_value_kernel(const OpParams P, int64_t count,
                          const double *__restrict__ tab, double *out)
    {
      const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
      if (i >= count)
        return;
      const double *r    = tab + i * (1 + dim);
      const double  x[3] = {r[1], r[2], dim == 3 ? r[3] : 0.};
      out[i] = analytic_solution<dim>(P.sol_kind, P.sol_param, P.sol_time,
                                      P.vel_const, x);
    }

    template <int dim, int n>
    void
    launch_geometry_tables(Operator &op, const OpParams &P)
    {
      const auto *B = op.d_basis_d.as<DevBasis<double>>();
      if (op.mesh.n_cells > 0)
        hipLaunchKernelGGL((geometry_table_kernel<dim, n>),
                           dim3(op.mesh.n_cells), dim3(aux_threads), 0,
                           op.stream, P, B, op.d_tab_cell[0].as<double>(),
                           op.d_tab_face[0].as<double>());
      hip_check(hipGetLastError(), "geometry tables");
    }

    template <int dim, int n>
    void
    launch_boundary_tables(Operator &op, const OpParams &P, const int *bcell,
                           const int *bface, int64_t n_bnd)
    {
      const auto *B = op.d_basis_d.as<DevBasis<double>>();
      if (n_bnd > 0)
        hipLaunchKernelGGL((boundary_table_kernel<dim, n>), dim3(n_bnd),
                           dim3(64), 0, op.stream, P, B, bcell, bface,
                           op.d_tab_bnd.as<double>());
      hip_check(hipGetLastError(), "boundary tables");
    }

    void
    reduce_to_host(Operator &op, int64_t count, int k, double *out)
    {
      op.d_result.alloc(sizeof(double) * 64);
      hipLaunchKernelGGL(reduce_columns, dim3(1), dim3(1024), 0, op.stream,
                         op.d_partials.as<double>(), count, k,
                         op.d_result.as<double>());
      hip_check(hipGetLastError(), "reduce");
      hip_check(hipMemcpyAsync(out, op.d_result.ptr, sizeof(double) * k,
                               hipMemcpyDeviceToHost, op.stream),
                "copy result");
      hip_check(hipStreamSynchronize(op.stream), "sync");
    }
  } // namespace

  void
  aux_project(const Operator &op, void *dst, double t)
  {
    const OpParams P = op.base_params(t, 1);
    const int      nc = op.cfg.n_components;
    if (op.mesh.n_cells == 0)
      return;
    if (op.cfg.dim == 3)
      hipLaunchKernelGGL(project_kernel<3>, dim3(op.mesh.n_cells),
                         dim3(aux_threads), 0, op.stream, P,
                         op.d_basis_d.as<DevBasis<double>>(), op.n, nc, t,
                         dst, (int)op.fp32());
    else
      hipLaunchKernelGGL(project_kernel<2>, dim3(op.mesh.n_cells),
                         dim3(aux_threads), 0, op.stream, P,
                         op.d_basis_d.as<DevBasis<double>>(), op.n, nc, t,
                         dst, (int)op.fp32());
    hip_check(hipGetLastError(), "project");
  }

  void
  aux_inverse_mass(const Operator &op, int p, void *vec)
  {
    const OpParams P = op.base_params(0, 1);
    if (op.mesh.n_cells == 0)
      return;
    const int fp32 = p == DG_FP32 ? 1 : 0;
    auto      it   = mass_registry().find(mass_key(op.cfg.dim, op.n, fp32));
    if (it == mass_registry().end())
      throw Error(DG_ERR_UNSUPPORTED, "inverse mass variant not built");
    hip_check(it->second(P, op.mass_basis_ptr(fp32 != 0),
                         op.mass_basis_ptr(false), vec, op.stream),
              "inverse mass");
  }

  void
  aux_boundary_tables(Operator &op)
  {
    if (op.boundary_ready)
      return;
    const int     dim = op.cfg.dim, n = op.n;
    const int     nl  = dim == 3 ? n * n : n;
    const int64_t nc  = op.mesh.n_cells;
    // boundary faces in cell order
    std::vector<int> bidx(std::max<int64_t>(1, nc * 2 * dim), -1);
    std::vector<int> bcell, bface;
    for (int64_t c = 0; c < nc; ++c)
      for (int f = 0; f < 2 * dim; ++f)
        if (op.mesh.neighbors[c * 2 * dim + f] < 0)
          {
            bidx[c * 2 * dim + f] = (int)bcell.size();
            bcell.push_back((int)c);
            bface.push_back(f);
          }
    const int64_t nb = (int64_t)bcell.size();
    op.d_bnd_index.alloc(sizeof(int) * bidx.size());
    hip_check(hipMemcpy(op.d_bnd_index.ptr, bidx.data(),
                        sizeof(int) * bidx.size(), hipMemcpyHostToDevice),
              "upload boundary index");
    op.d_tab_bnd.alloc(sizeof(double) * (1 + dim) * nl *
                       std::max<int64_t>(1, nb));
    op.d_bnd_value.alloc(sizeof(double) * nl * std::max<int64_t>(1, nb));
    op.n_bnd = nb;
    if (nb > 0)
      {
        DevBuf d_bcell, d_bface;
        d_bcell.alloc(sizeof(int) * nb);
        d_bface.alloc(sizeof(int) * nb);
        hip_check(hipMemcpy(d_bcell.ptr, bcell.data(), sizeof(int) * nb,
                            hipMemcpyHostToDevice),
                  "upload boundary cells");
        hip_check(hipMemcpy(d_bface.ptr, bface.data(), sizeof(int) * nb,
                            hipMemcpyHostToDevice),
                  "upload boundary faces");
        const OpParams P = op.base_params(0, 1);
#define DG_BT(D, N)                                                        \
  else if (dim == D && n == N)                                             \
    launch_boundary_tables<D, N>(op, P, d_bcell.as<int>(), d_bface.as<int>(), \
                                 nb);
        if (false)
          ;
        DG_BT(2, 2) DG_BT(2, 3) DG_BT(2, 4) DG_BT(2, 5) DG_BT(2, 6)
        DG_BT(2, 7) DG_BT(2, 8) DG_BT(3, 2) DG_BT(3, 3) DG_BT(3, 4)
        DG_BT(3, 5) DG_BT(3, 6) DG_BT(3, 7) DG_BT(3, 8)
#undef DG_BT
        // the temporaries above are freed at scope end: finish first
        hip_check(hipStreamSynchronize(op.stream), "boundary tables");
      }
    op.boundary_ready = true;
  }

  void
  aux_geometry_tables(Operator &op, int p)
  {
    const int pi = p == DG_FP32 ? 1 : 0;
    if (op.tables_ready[pi])
      return;
    aux_boundary_tables(op);
    const int     dim = op.cfg.dim, n = op.n;
    const int     nl  = dim == 3 ? n * n : n;
    const int64_t nc  = op.mesh.n_cells;
    if (!op.tables_ready[0])
      {
        op.d_tab_cell[0].alloc(sizeof(double) * (dim + 1) * op.n_dofs);
        op.d_tab_face[0].alloc(sizeof(double) * dim * nl *
                               std::max<int64_t>(1, nc));
        const OpParams P = op.base_params(0, 1);
#define DG_GT(D, N)                                                         \
  else if (dim == D && n == N) launch_geometry_tables<D, N>(op, P);
        if (false)
          ;
        DG_GT(2, 2) DG_GT(2, 3) DG_GT(2, 4) DG_GT(2, 5) DG_GT(2, 6)
        DG_GT(2, 7) DG_GT(2, 8) DG_GT(3, 2) DG_GT(3, 3) DG_GT(3, 4)
        DG_GT(3, 5) DG_GT(3, 6) DG_GT(3, 7) DG_GT(3, 8)
#undef DG_GT
        op.tables_ready[0] = true;
      }
    if (pi == 1)
      {
        const int64_t ncell = (dim + 1) * op.n_dofs;
        const int64_t nface = (int64_t)dim * nl * nc;
        op.d_tab_cell[1].alloc(sizeof(float) * std::max<int64_t>(1, ncell));
        op.d_tab_face[1].alloc(sizeof(float) * std::max<int64_t>(1, nface));
        vec_copy_convert(op, ncell, DG_FP64, op.d_tab_cell[0].ptr, DG_FP32,
                         op.d_tab_cell[1].ptr);
        vec_copy_convert(op, nface, DG_FP64, op.d_tab_face[0].ptr, DG_FP32,
                         op.d_tab_face[1].ptr);
        op.tables_ready[1] = true;
      }
  }

  void
  aux_boundary_values(Operator &op, const OpParams &P)
  {
    if (op.n_bnd == 0)
      return;
    const int     nl    = op.cfg.dim == 3 ? op.n * op.n : op.n;
    const int64_t count = op.n_bnd * nl;
    const unsigned nb   = (unsigned)((count + aux_threads - 1) / aux_threads);
    if (op.cfg.dim == 3)
      hipLaunchKernelGGL(boundary_value_kernel<3>, dim3(nb), dim3(aux_threads),
                         0, op.stream, P, count, op.d_tab_bnd.as<double>(),
                         op.d_bnd_value.as<double>());
    else
      hipLaunchKernelGGL(boundary_value_kernel<2>, dim3(nb), dim3(aux_threads),
                         0, op.stream, P, count, op.d_tab_bnd.as<double>(),
                         op.d_bnd_value.as<double>());
    hip_check(hipGetLastError(), "boundary values");
  }

  void
  aux_fdm_cell_data(const Operator &op, double *velocity, double *invdet)
  {
    const OpParams P = op.base_params(0, 1);
    if (op.mesh.n_cells == 0)
      return;
    if (op.cfg.dim == 3)
      hipLaunchKernelGGL(fdm_cell_data_kernel<3>, dim3(op.mesh.n_cells),
                         dim3(aux_threads), 0, op.stream, P,
                         op.d_basis_d.as<DevBasis<double>>(), op.n, velocity,
                         invdet);
    else
      hipLaunchKernelGGL(fdm_cell_data_kernel<2>, dim3(op.mesh.n_cells),
                         dim3(aux_threads), 0, op.stream, P,
                         op.d_basis_d.as<DevBasis<double>>(), op.n, velocity,
                         invdet);
    hip_check(hipGetLastError(), "fdm cell data");
  }

  void
  aux_mass_energy(const Operator &op_c, const void *vec, double out[3])
  {
    Operator &op = const_cast<Operator &>(op_c);
    op.d_partials.alloc(sizeof(double) * 3 * std::max<int64_t>(1, op.mesh.n_cells));
    const OpParams P = op.base_params(0, 1);
    if (op.cfg.dim == 3)
      hipLaunchKernelGGL(mass_energy_kernel<3>, dim3(op.mesh.n_cells),
                         dim3(aux_threads), 0, op.stream, P,
                         op.d_basis_d.as<DevBasis<double>>(), op.n, vec,
                         (int)op.fp32(), op.d_partials.as<double>());
    else
      hipLaunchKernelGGL(mass_energy_kernel<2>, dim3(op.mesh.n_cells),
                         dim3(aux_threads), 0, op.stream, P,
                         op.d_basis_d.as<DevBasis<double>>(), op.n, vec,
                         (int)op.fp32(), op.d_partials.as<double>());
    hip_check(hipGetLastError(), "mass_energy");
    reduce_to_host(op, op.mesh.n_cells, 3, out);
  }

  void
  aux_l2_error(const Operator &op_c, const void *vec, double t, double out[2])
  {
    Operator &op = const_cast<Operator &>(op_c);
    op.d_partials.alloc(sizeof(double) * 2 * std::max<int64_t>(1, op.mesh.n_cells));
    const OpParams   P = op.base_params(t, 1);
    const QuadTables Q = make_quad_tables(op.n);
    if (op.cfg.dim == 3)
      hipLaunchKernelGGL(l2_error_kernel<3>, dim3(op.mesh.n_cells),
                         dim3(aux_threads), 0, op.stream, P,
                         op.d_basis_d.as<DevBasis<double>>(), Q, op.n, vec,
                         (int)op.fp32(), t, op.d_partials.as<double>());
    else
      hipLaunchKernelGGL(l2_error_kernel<2>, dim3(op.mesh.n_cells),
                         dim3(aux_threads), 0, op.stream, P,
                         op.d_basis_d.as<DevBasis<double>>(), Q, op.n, vec,
                         (int)op.fp32(), t, op.d_partials.as<double>());
    hip_check(hipGetLastError(), "l2_error");
    double s[2];
    reduce_to_host(op, op.mesh.n_cells, 2, s);
    out[0] = std::sqrt(s[0]);
    out[1] = std::sqrt(s[1]);
  }

  void
  vec_axpby(Operator &op, int p, int64_t n, double a, const void *x, double b,
            void *y)
  {
    if (n <= 0)
      return;
    if (p == DG_FP32)
      hipLaunchKernelGGL(axpby_kernel<float>, dim3(grid_for(n)),
                         dim3(vec_threads), 0, op.stream, n, (float)a,
                         (const float *)x, (float)b, (float *)y);
    else
      hipLaunchKernelGGL(axpby_kernel<double>, dim3(grid_for(n)),
                         dim3(vec_threads), 0, op.stream, n, a,
                         (const double *)x, b, (double *)y);
    hip_check(hipGetLastError(), "axpby");
  }

  void
  vec_multi_axpy(Operator &op, int p, int64_t n, int m, const double *alpha,
                 const void *x, int64_t xs, void *y)
  {
    if (n <= 0 || m <= 0)
      return;
    if (p == DG_FP32)
      multi_axpy_t<float>(op, n, m, alpha, (const float *)x, xs, (float *)y);
    else
      multi_axpy_t<double>(op, n, m, alpha, (const double *)x, xs,
                           (double *)y);
  }

  void
  vec_multi_axpy_scaled(Operator &op, int p, int64_t n, int m,
                        const double *alpha, const void *x, int64_t xs,
                        void *y, double scale)
  {
    if (n <= 0 || m <= 0)
      return;
    if (p == DG_FP32)
      multi_axpy_t<float>(op, n, m, alpha, (const float *)x, xs, (float *)y,
                          scale);
    else
      multi_axpy_t<double>(op, n, m, alpha, (const double *)x, xs,
                           (double *)y, scale);
  }

  void
  vec_multi_axpy_dot(Operator &op, int p, int64_t n, int m,
                     const double *alpha, const void *x, int64_t xs, void *y,
                     double *result)
  {
    if (m <= 0)
      return;
    if (m > max_fused)
      {
        // more vectors than one fused pass holds: two passes
        vec_multi_axpy(op, p, n, m, alpha, x, xs, y);
        vec_multi_dot(op, p, n, m, x, xs, y, result);
        vec_multi_dot(op, p, n, 1, y, n, y, result + m);
        return;
      }
    if (n <= 0)
      {
        for (int j = 0; j <= m; ++j)
          result[j] = 0;
        return;
      }
    if (p == DG_FP32)
      multi_axpy_dot_t<float>(op, n, m, alpha, (const float *)x, xs,
                              (float *)y, result);
    else
      multi_axpy_dot_t<double>(op, n, m, alpha, (const double *)x, xs,
                               (double *)y, result);
  }

  void
  vec_multi_dot(Operator &op, int p, int64_t n, int m, const void *x,
                int64_t xs, const void *y, double *result)
  {
    if (m <= 0)
      return;
    if (n <= 0)
      {
        for (int j = 0; j < m; ++j)
          result[j] = 0;
        return;
      }
    if (p == DG_FP32)
      multi_dot_t<float>(op, n, m, (const float *)x, xs, (const float *)y,
                         result);
    else
      multi_dot_t<double>(op, n, m, (const double *)x, xs, (const double *)y,
                          result);
  }

  void
  vec_multi_dot_dev(Operator &op, int p, int64_t n, int m, const void *x,
                    int64_t xs, const void *y, double *d_out)
  {
    if (m <= 0)
      return;
    if (n <= 0)
      throw Error(DG_ERR_ARGUMENT, "multi_dot_dev: empty vectors");
    if (p == DG_FP32)
      multi_dot_dev_t<float>(op, n, m, (const float *)x, xs,
                             (const float *)y, d_out);
    else
      multi_dot_dev_t<double>(op, n, m, (const double *)x, xs,
                              (const double *)y, d_out);
  }

  // ---- multi-dot without allocation ---------------------------------------
  //
  // dg_vector_multi_dot_noalloc keeps its block partials (one chunk of up to
  // max_multi vectors over at most noalloc_blocks workgroups) and the m sums
  // in this array, one copy per GPU for the life of the process, instead of
  // the operator's allocated buffers. In host code `scratch` names no
  // device address: kernels get the address hipGetSymbolAddress returns for
  // the device (noalloc_scratch below). Every call on one GPU uses the same
  // array, so calls that overlap on one GPU (two streams, two threads)
  // race. Nothing in it has to start at zero: every partial and sum is
  // written before it is read.
  namespace
  {
    constexpr int noalloc_blocks  = 1024;
    constexpr int noalloc_max_vec = 64;
    constexpr int noalloc_doubles = noalloc_blocks * max_multi + noalloc_max_vec;

    static __device__ __align__(256) double scratch[noalloc_doubles];

    // the device address of `scratch` on `device`, looked up once per
    // device with that device current
    double *
    noalloc_scratch(int device)
    {
      static std::mutex              mu;
      static std::map<int, double *> ptrs;
      std::lock_guard<std::mutex>    lock(mu);
      auto                           it = ptrs.find(device);
      if (it != ptrs.end())
        return it->second;
      int cur = 0;
      hip_check(hipGetDevice(&cur), "hipGetDevice");
      if (cur != device)
        hip_check(hipSetDevice(device), "hipSetDevice");
      void            *p  = nullptr;
      const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (cur != device)
        hip_check(hipSetDevice(cur), "hipSetDevice");
      hip_check(rc, "hipGetSymbolAddress(scratch)");
      if (p == nullptr)
        throw Error(DG_ERR_HIP, "hipGetSymbolAddress(scratch): null address");
      ptrs[device] = static_cast<double *>(p);
      return ptrs[device];
    }

    template <typename T>
    void
    multi_dot_noalloc_t(Operator &op, double *scr, int64_t n, int m,
                        const T *x, int64_t xs, const T *y, double *result)
    {
      const unsigned nb  = std::min<unsigned>(grid_for(n), noalloc_blocks);
      double        *part = scr;
      double        *sums = scr + noalloc_blocks * max_multi;
      int            done = 0;
      while (done < m)
        {
          const int chunk = std::min(max_multi, m - done);
          const T  *xc    = x + done * xs;
#define DG_MD(M)                                                              \
  case M:                                                                     \
    hipLaunchKernelGGL((multi_dot_kernel<T, M>), dim3(nb), dim3(vec_threads), \
                       0, op.stream, n, xc, xs, y, part);                     \
    break;
          switch (chunk)
            {
              DG_MD(1) DG_MD(2) DG_MD(3) DG_MD(4) DG_MD(5) DG_MD(6) DG_MD(7)
              DG_MD(8) DG_MD(9) DG_MD(10) DG_MD(11) DG_MD(12) DG_MD(13)
              DG_MD(14) DG_MD(15) DG_MD(16)
            }
#undef DG_MD
          hip_check(hipGetLastError(), "multi_dot");
          hipLaunchKernelGGL(reduce_columns_small, dim3(1), dim3(1024), 0,
                             op.stream, part, (int)nb, chunk, sums + done);
          hip_check(hipGetLastError(), "reduce");
          done += chunk;
        }
      hip_check(hipMemcpyAsync(result, sums, sizeof(double) * m,
                               hipMemcpyDeviceToHost, op.stream),
                "copy result");
      hip_check(hipStreamSynchronize(op.stream), "sync");
    }
  } // namespace

  void
  vec_multi_dot_noalloc(Operator &op, int p, int64_t n, int m, const void *x,
                        int64_t xs, const void *y, double *result)
  {
    if (m > noalloc_max_vec)
      throw Error(DG_ERR_ARGUMENT,
                  "multi_dot_noalloc: at most " +
                    std::to_string(noalloc_max_vec) + " vectors, got " +
                    std::to_string(m));
    if (m <= 0)
      return;
    if (n <= 0)
      {
        for (int j = 0; j < m; ++j)
          result[j] = 0;
        return;
      }
    double *scr = noalloc_scratch(op.device);
    if (p == DG_FP32)
      multi_dot_noalloc_t<float>(op, scr, n, m, (const float *)x, xs,
                                 (const float *)y, result);
    else
      multi_dot_noalloc_t<double>(op, scr, n, m, (const double *)x, xs,
                                  (const double *)y, result);
  }

  void
  vec_scale_inv_sqrt(Operator &op, int p, int64_t n, const double *d_s2,
                     void *x)
  {
    if (n <= 0)
      return;
    if (p == DG_FP32)
      hipLaunchKernelGGL(scale_inv_sqrt_kernel<float>, dim3(grid_for(n)),
                         dim3(vec_threads), 0, op.stream, n, d_s2,
                         (float *)x);
    else
      hipLaunchKernelGGL(scale_inv_sqrt_kernel<double>, dim3(grid_for(n)),
                         dim3(vec_threads), 0, op.stream, n, d_s2,
                         (double *)x);
    hip_check(hipGetLastError(), "scale");
  }

  void
  vec_arnoldi_update(Operator &op, int p, int64_t n, int m, const double *d_h,
                     const void *V, int64_t vs, void *w)
  {
    if (m < 1 || m > max_arnoldi)
      throw Error(DG_ERR_ARGUMENT, "arnoldi update: 1..16 basis vectors");
    if (n <= 0)
      return;
    if (p == DG_FP32)
      arnoldi_update_t<float>(op, n, m, d_h, (const float *)V, vs,
                              (float *)w);
    else
      arnoldi_update_t<double>(op, n, m, d_h, (const double *)V, vs,
                               (double *)w);
  }

  void
  vec_arnoldi_lsq(Operator &op, int m, const double *d_h, int64_t hs,
                  const double *d_beta2, double *d_y)
  {
    if (m < 1 || m > max_arnoldi)
      throw Error(DG_ERR_ARGUMENT, "arnoldi least squares: 1..16 columns");
    hipLaunchKernelGGL(arnoldi_lsq_kernel, dim3(1), dim3(64), 0, op.stream, m,
                       d_h, hs, d_beta2, d_y);
    hip_check(hipGetLastError(), "arnoldi least squares");
  }

  void
  vec_combine(Operator &op, int p, int64_t n, int m, const double *d_y,
              const void *Z, int64_t zs, double *x)
  {
    if (m < 1 || m > max_arnoldi)
      throw Error(DG_ERR_ARGUMENT, "combine: 1..16 vectors");
    if (n <= 0)
      return;
    if (p == DG_FP32)
      combine_t<float>(op, n, m, d_y, (const float *)Z, zs, x);
    else
      combine_t<double>(op, n, m, d_y, (const double *)Z, zs, x);
  }

  void
  vec_gram4(Operator &op, int64_t n, const double *const v[4],
            const double *rhs, double out[14])
  {
    const unsigned nb = std::min<unsigned>(grid_for(n), 1024);
    op.d_partials.alloc(sizeof(double) * nb * 14);
    hipLaunchKernelGGL(gram4_kernel, dim3(nb), dim3(vec_threads), 0,
                       op.stream, n, v[0], v[1], v[2], v[3], rhs,
                       op.d_partials.as<double>());
    hip_check(hipGetLastError(), "gram4");
    reduce_to_host(op, nb, 14, out);
  }

  void
  vec_copy_convert(Operator &op, int64_t n, int ps, const void *src, int pd,
                   void *dst)
  {
    if (n <= 0)
      return;
    const dim3 g(grid_for(n)), b(vec_threads);
    if (ps == DG_FP64 && pd == DG_FP64)
      hipLaunchKernelGGL((convert_kernel<double, double>), g, b, 0, op.stream,
                         n, (const double *)src, (double *)dst);
    else if (ps == DG_FP64 && pd == DG_FP32)
      hipLaunchKernelGGL((convert_kernel<double, float>), g, b, 0, op.stream,
                         n, (const double *)src, (float *)dst);
    else if (ps == DG_FP32 && pd == DG_FP64)
      hipLaunchKernelGGL((convert_kernel<float, double>), g, b, 0, op.stream,
                         n, (const float *)src, (double *)dst);
    else
      hipLaunchKernelGGL((convert_kernel<float, float>), g, b, 0, op.stream,
                         n, (const float *)src, (float *)dst);
    hip_check(hipGetLastError(), "convert");
  }

  void
  vec_lsrk_update(Operator &op, int p, int64_t n, double b, double a,
                  const void *k, const void *sol_in, void *sol_out,
                  void *ri_out)
  {
    if (n <= 0)
      return;
    if (p == DG_FP32)
      hipLaunchKernelGGL(lsrk_update_kernel<float>, dim3(grid_for(n)),
                         dim3(vec_threads), 0, op.stream, n, (float)b,
                         (float)a, (const float *)k, (const float *)sol_in,
                         (float *)sol_out, (float *)ri_out);
    else
      hipLaunchKernelGGL(lsrk_update_kernel<double>, dim3(grid_for(n)),
                         dim3(vec_threads), 0, op.stream, n, b, a,
                         (const double *)k, (const double *)sol_in,
                         (double *)sol_out, (double *)ri_out);
    hip_check(hipGetLastError(), "lsrk_update");
  }
} // namespace dgops
