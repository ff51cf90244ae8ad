// Inviscid Burgers right-hand side M^-1(-L(u)) of a 3-component field in 3D
// for degrees 1..7 (burgers_solver.cc:259-341, the degree a run-time
// argument there, :936-938): cell flux u (x) u and the Rusanov face flux
//   F = 1/2 [(u-.n) u- + (u+.n) u+ + alpha max(|u-.n|, |u+.n|)(u- - u+)]
// integrated with QGauss(m), m = k(k+1)/2 (burgers_solver.cc:234), and the
// inverse mass matrix of QGauss(k+1) (:235, :366).
//
// The m^3 quadrature tensors (28^3 points at degree 7) do not fit in LDS,
// so the cell term runs slab by slab: the nodal values are first contracted
// along z to the m quadrature levels (W), and for every level qz the slab
// is interpolated in x and y, the fluxes u_c v_d (v = w adj(J) u) are formed
// point-wise, integrated in x and y, and the z integration is accumulated
// in registers of the threads that own the cell's z-lines:
//   res_c = sum_qz S_z[qz] (x) (S_y^T G_x^T f_cx + G_y^T S_x^T f_cy)
//         + G_z[qz] (x) S_y^T S_x^T f_cz.
// Faces are element-centric (own outward normal, neighbour trace from the
// neighbour's nodal face layer) and are integrated first into an LDS
// accumulator that seeds the registers. CPB cells per workgroup, R threads
// per cell; every phase loops its tasks over the cell's R threads.
#pragma once

#include "burgers_kernel.hpp"

namespace dgops
{
  constexpr int
  burgers3_threads(int n, int m)
  {
    const int a = m * m, b = n * m, c = n * n;
    return a > b ? (a > c ? a : c) : (b > c ? b : c);
  }

  // cells per workgroup: up to 128 threads of cells
  constexpr int
  burgers3_cells(int n, int m)
  {
    const int r = burgers3_threads(n, m);
    return r >= 128 ? 1 : 128 / r;
  }

  template <int f>
  struct FaceId
  {
    static constexpr int value = f;
  };

  // face f = 2 d + s of a cell: its normal direction d, side s, tangential
  // directions ta < tb, and node (a, b) of the layer at `layer` along d in
  // the lexicographic cell numbering (x fastest)
  template <int n, int f>
  struct FaceIdx
  {
    static constexpr int d = f / 2, s = f % 2;
    static constexpr int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    static constexpr int sd = d == 0 ? 1 : d == 1 ? n : n * n;
    static constexpr int sa = ta == 0 ? 1 : ta == 1 ? n : n * n;
    static constexpr int sb = tb == 0 ? 1 : tb == 1 ? n : n * n;
    __device__ static constexpr int
    node(int a, int b, int layer)
    {
      return layer * sd + a * sa + b * sb;
    }
  };

  template <int n, int m>
  struct Burgers3Lay
  {
    static constexpr int n2 = n * n, n3 = n * n * n, m2 = m * m;
    static constexpr int R  = burgers3_threads(n, m);
    // per-cell LDS offsets (doubles)
    static constexpr int oU  = 0;                // [3][n3] nodal input
    static constexpr int oW  = oU + 3 * n3;      // [2][3][n2] two levels
    static constexpr int oT  = oW + 6 * n2;      // [3][n][m]
    static constexpr int oUV = oT + 3 * n * m;   // [6][m2]
    static constexpr int oF  = oUV + 6 * m2;     // [9][m][n]
    static constexpr int oX  = oF + 9 * m * n;   // [8][3] corners
    static constexpr int size = oX + 24;
    // unroll of the loops over the m points of a line
    static constexpr int UM = m <= 12 ? m : 4;
  };

  // four wavefronts per SIMD where that costs no spills (m <= 6, degree
  // <= 3: 128 VGPRs with 20 KB of LDS per 3-cell workgroup at degree 3)
  template <int n, int m, bool cartesian, int epi, int CPB>
  __global__ void __launch_bounds__(CPB * burgers3_threads(n, m))
  __attribute__((amdgpu_waves_per_eu(m <= 6 ? 4 : 1)))
  burgers3_kernel(const OpParams P, const DevBasis<double> *__restrict__ Bd,
                  const BurgersTables *__restrict__ BT,
                  const OpBuffers<double> io)
  {
    using L             = Burgers3Lay<n, m>;
    constexpr int n2    = L::n2, n3 = L::n3, m2 = L::m2, R = L::R;
    constexpr int nvec  = 3 * n3;
    __shared__ double lds[CPB * L::size];
    __shared__ double sS[m * n], sG[m * n], sw[m], sx[m];

    const int     tid    = threadIdx.x;
    const int     lc     = tid / R;
    const int     r      = tid - lc * R;
    const int64_t cell   = (int64_t)xcd_block(blockIdx.x, gridDim.x) * CPB +
                         lc;
    const bool    active = cell < P.n_cells;
    double       *U  = lds + lc * L::size + L::oU;
    double       *W  = lds + lc * L::size + L::oW;
    double       *T  = lds + lc * L::size + L::oT;
    double       *UV = lds + lc * L::size + L::oUV;
    double       *F  = lds + lc * L::size + L::oF;
    double       *X  = lds + lc * L::size + L::oX;
    const double *src = io.src + cell * nvec;

    // ---- inputs: nodal values, corners, the QGauss(m) tables -----------
    for (int e = tid; e < m * n; e += blockDim.x)
      {
        sS[e] = BT->Sm[e / n][e % n];
        sG[e] = BT->Gm[e / n][e % n];
      }
    for (int e = tid; e < m; e += blockDim.x)
      {
        sw[e] = BT->wm[e];
        sx[e] = BT->xm[e];
      }
    for (int e = r; e < nvec; e += R)
      U[e] = active ? src[e] : 0.;
    if (!cartesian)
      for (int e = r; e < 24; e += R)
        X[e] = active ? P.cell_vertices[cell * 24 + e] : 0.;
    __syncthreads();

    // ---- faces: own and neighbour traces at the m x m face points, the
    // Rusanov flux with this cell's outward normal, integrated onto the own
    // face layer. Both faces of a direction at once (their layers are
    // disjoint); one instance per direction, so the strides are static. The
    // face buffers overlay the cell-term buffers W..F, unused until the
    // cell term: FT [side s][own/neighbour][c][b m + qa] (12 n m values),
    // FL [s][c][qb m + qa], the flux on the own side (6 m^2 values)
    static_assert(12 * n * m + 6 * m2 <= L::oX - L::oW,
                  "face buffers overlay the cell-term buffers");
    double *FT  = W;
    double *FL  = W + 12 * n * m;
    // the x and y faces' contributions to their layers, [s][c][b n + a],
    // in the rest of the overlay (6 n^2 values: what W..F leave after FT and
    // FL), integrated by all n^2 threads and added by the layer's owners
    // after the next barrier
    double *FV = FL + 6 * m2;
    static_assert(12 * n * m + 6 * m2 + 6 * n2 <= L::oX - L::oW,
                  "face layer buffer fits the overlay");
    // z-line owners (thread r < n2: i = r % n, j = r / n) accumulate the
    // face and cell terms of their line in registers
    double acc[3][n];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int k = 0; k < n; ++k)
        acc[c][k] = 0;
    // per-thread tasks, fixed for every phase below: each phase has at most
    // R tasks (R >= n m, m^2, n^2), one per thread, and a task covers all
    // components and sides
    const bool    t_nm = r < n * m, t_m2 = r < m2, t_n2 = r < n2;
    const int     xj = r / m, xq = r % m; // (j or b, qx or qa)
    const int     yq = r % m, yy = r / m; // (qa, qb) or (qx, qy)
    const int     gi = r % n, gq = r / n; // (i, qy)
    const int     hi = r % n, hj = r / n; // (i, j) or (a, b)
    // the owners of the layers of direction dp (0: x, 1: y) add FV
    auto take_layers = [&](auto dc) {
      constexpr int dp = decltype(dc)::value;
      if (t_n2)
        {
          const int i = hi, j = hj;
#pragma unroll
          for (int s = 0; s < 2; ++s)
            {
              const int layer = s ? n - 1 : 0;
              if (dp == 0 ? i == layer : j == layer)
                {
                  const int a = dp == 0 ? j : i;
#pragma unroll
                  for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int b = 0; b < n; ++b)
                      acc[c][b] += FV[(s * 3 + c) * n2 + b * n + a];
                }
            }
        }
    };
    auto    face_pair = [&](auto dc) {
      constexpr int d = decltype(dc)::value;
      using FI        = FaceIdx<n, 2 * d>; // strides of direction d
      constexpr int ta = FI::ta, tb = FI::tb;
      if constexpr (d > 0)
        take_layers(FaceId<d - 1>{});
      // along ta: FT[(s 2 + side) 3 + c][b m + qa], thread (b, qa)
      if (t_nm)
        {
          const int b = xj, qa = xq;
          // row b of the neighbours' layers facing this cell (face s: the
          // neighbour's layer 1 - s), from global memory; across a boundary
          // face the row is zero (the load reads the cell's own row, or
          // cell 0's for a thread without a cell, and is discarded)
          double nv[2][3][n];
#pragma unroll
          for (int s = 0; s < 2; ++s)
            {
              const int64_t nb =
                active ? (int64_t)P.neighbors[cell * 6 + 2 * d + s] : -1;
              const bool    has = nb >= 0;
              const int64_t nbc = has ? nb : active ? cell : 0;
#pragma unroll
              for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int a = 0; a < n; ++a)
                  {
                    const double v = io.src[nbc * nvec + c * n3 +
                                            FI::node(a, b, s ? 0 : n - 1)];
                    nv[s][c][a]    = has ? v : 0.;
                  }
            }
#pragma unroll
          for (int k = 0; k < 12; ++k)
            {
              const int s = k / 6, sd = (k / 3) % 2, c = k % 3;
              double    v = 0;
#pragma unroll
              for (int a = 0; a < n; ++a)
                {
                  const double x =
                    sd == 0 ? U[c * n3 + FI::node(a, b, s ? n - 1 : 0)] :
                              nv[s][c][a];
                  v = fma(sS[qa * n + a], x, v);
                }
              FT[k * n * m + b * m + qa] = v;
            }
        }
      __syncthreads();
      // along tb, then the flux, at the face point of thread (qa, qb),
      // both sides; FL receives -F n dA of the own side
      if (t_m2)
        {
          const int qa = yq, qb = yy, pt = r;
          double    tr[12]; // [s][own/neighbour][c]
#pragma unroll
          for (int k = 0; k < 12; ++k)
            {
              double v = 0;
#pragma unroll
              for (int b = 0; b < n; ++b)
                v = fma(sS[qb * n + b], FT[k * n * m + b * m + qa], v);
              tr[k] = v;
            }
          const double w = sw[qa] * sw[qb];
#pragma unroll
          for (int s = 0; s < 2; ++s)
            {
              // normal velocities u.n on both sides with the unit outward
              // normal n, and the area element len
              const double *um = tr + 6 * s, *up = um + 3;
              double        len, unm = 0, unp = 0;
              if constexpr (cartesian)
                {
                  // n = -+e_d, area element h_ta h_tb w
                  len = P.h[ta] * P.h[tb] * w;
                  unm = s ? um[d] : -um[d];
                  unp = s ? up[d] : -up[d];
                }
              else
                {
                  double xi[3];
                  xi[d]  = s;
                  xi[ta] = sx[qa];
                  xi[tb] = sx[qb];
                  double x[3], J[3][3], adj[3][3], nrm[3];
                  multilinear_map<3>(X, xi, x, J);
                  adjugate<3>(J, adj);
                  for (int e = 0; e < 3; ++e)
                    nrm[e] = (s ? 1. : -1.) * adj[d][e] * w;
                  len = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] +
                             nrm[2] * nrm[2]);
                  for (int c = 0; c < 3; ++c)
                    {
                      unm = fma(um[c], nrm[c], unm);
                      unp = fma(up[c], nrm[c], unp);
                    }
                  const double il = 1. / len;
                  unm *= il;
                  unp *= il;
                }
              const double lam = P.alpha * fmax(fabs(unm), fabs(unp));
              for (int c = 0; c < 3; ++c)
                FL[(s * 3 + c) * m2 + pt] =
                  -0.5 * (unm * um[c] + unp * up[c] + lam * (um[c] - up[c])) *
                  len;
            }
        }
      __syncthreads();
      // integrate: along tb (into FT[s 3 + c]), thread (b, qa) ...
      if (t_nm)
        {
          const int b = xj, qa = xq;
#pragma unroll
          for (int k = 0; k < 6; ++k)
            {
              const double *fl = FL + k * m2 + qa;
              double        v  = 0;
#pragma unroll L::UM
              for (int qb = 0; qb < m; ++qb)
                v = fma(sS[qb * n + b], fl[qb * m], v);
              FT[k * n * m + b * m + qa] = v;
            }
        }
      __syncthreads();
      // ... then along ta onto the own layer of side s: z faces by every
      // z-line owner (i, j) into its register of node k = layer; x and y
      // faces by thread (a, b) into FV (node (layer, a, b) of the x faces,
      // (a, layer, b) of the y faces), added by the owners (layer, j) or
      // (i, layer) of those lines after the barrier
      if (t_n2)
        {
          const int i = hi, j = hj;
#pragma unroll
          for (int s = 0; s < 2; ++s)
            {
              const int layer = s ? n - 1 : 0;
              if constexpr (d == 2)
                {
#pragma unroll
                  for (int c = 0; c < 3; ++c)
                    {
                      const double *ft = FT + (s * 3 + c) * n * m + j * m;
                      double        v  = 0;
#pragma unroll L::UM
                      for (int qa = 0; qa < m; ++qa)
                        v = fma(sS[qa * n + i], ft[qa], v);
                      acc[c][layer] += v;
                    }
                }
              else
                {
                  const int a = i, b = j;
#pragma unroll
                  for (int c = 0; c < 3; ++c)
                    {
                      const double *ft = FT + (s * 3 + c) * n * m + b * m;
                      double        v  = 0;
#pragma unroll L::UM
                      for (int qa = 0; qa < m; ++qa)
                        v = fma(sS[qa * n + a], ft[qa], v);
                      FV[(s * 3 + c) * n2 + b * n + a] = v;
                    }
                }
            }
        }
      __syncthreads();
    };
    face_pair(FaceId<0>{});
    face_pair(FaceId<1>{});
    face_pair(FaceId<2>{});

    // ---- cell term, slab by slab ----------------------------------------
    // level qz of W[qz % 2][c][j n + i] = sum_k S[qz][k] u[c][k][j][i],
    // formed by the z-line owners from their nodal line (in U) one level
    // ahead, during the y phase of the level before
    auto w_level = [&](int qz) {
      double *w = W + (qz % 2) * 3 * n2 + r;
#pragma unroll
      for (int c = 0; c < 3; ++c)
        {
          double v = 0;
#pragma unroll
          for (int k = 0; k < n; ++k)
            v = fma(sS[qz * n + k], U[c * n3 + k * n2 + r], v);
          w[c * n2] = v;
        }
    };
    if (t_n2)
      w_level(0);
    __syncthreads();
    double vol = 0, adiag[3] = {0, 0, 0};
    if constexpr (cartesian)
      {
        vol = P.h[0] * P.h[1] * P.h[2];
        for (int e = 0; e < 3; ++e)
          adiag[e] = vol / P.h[e];
      }
    // the slab phases use the per-thread tasks of the face phases: x
    // interpolation (j, qx), y interpolation (qx, qy), x integration
    // (i, qy), y integration (i, j)
#pragma unroll 1
    for (int qz = 0; qz < m; ++qz)
      {
        // along x: T[c][j m + qx]
        if (t_nm)
          {
            const double *w = W + (qz % 2) * 3 * n2 + xj * n;
#pragma unroll
            for (int c = 0; c < 3; ++c)
              {
                double v = 0;
#pragma unroll
                for (int i = 0; i < n; ++i)
                  v = fma(sS[xq * n + i], w[c * n2 + i], v);
                T[c * n * m + xj * m + xq] = v;
              }
          }
        __syncthreads();
        // the next level of W (its buffer was last read by the x phase of
        // the level before)
        if (t_n2 && qz + 1 < m)
          w_level(qz + 1);
        // along y, then the point-wise fluxes: UV = (u_0, u_1, u_2, v_0,
        // v_1, v_2) with v = w adj(J) u
        if (t_m2)
          {
            double u[3];
#pragma unroll
            for (int c = 0; c < 3; ++c)
              {
                double v = 0;
#pragma unroll
                for (int j = 0; j < n; ++j)
                  v = fma(sS[yy * n + j], T[c * n * m + j * m + yq], v);
                u[c] = v;
              }
            const double w = sw[yq] * sw[yy] * sw[qz];
            double       v[3];
            if constexpr (cartesian)
              {
                for (int e = 0; e < 3; ++e)
                  v[e] = w * adiag[e] * u[e];
              }
            else
              {
                const double xi[3] = {sx[yq], sx[yy], sx[qz]};
                double       x[3], J[3][3], adj[3][3];
                multilinear_map<3>(X, xi, x, J);
                adjugate<3>(J, adj);
                for (int e = 0; e < 3; ++e)
                  v[e] = w * (adj[e][0] * u[0] + adj[e][1] * u[1] +
                              adj[e][2] * u[2]);
              }
            for (int c = 0; c < 3; ++c)
              {
                UV[c * m2 + r]       = u[c];
                UV[(3 + c) * m2 + r] = v[c];
              }
          }
        __syncthreads();
        // integrate along x: F[c 3 + branch][qy n + i], branches d = x (G),
        // y (S), z (S)
        if (t_nm)
          {
            double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            const double *uv = UV + gq * m;
#pragma unroll L::UM
            for (int qx = 0; qx < m; ++qx)
              {
                const double g  = sG[qx * n + gi], sv = sS[qx * n + gi];
                const double c0 = g * uv[3 * m2 + qx];
                const double c1 = sv * uv[4 * m2 + qx];
                const double c2 = sv * uv[5 * m2 + qx];
#pragma unroll
                for (int c = 0; c < 3; ++c)
                  {
                    const double uc = uv[c * m2 + qx];
                    a[c][0]         = fma(c0, uc, a[c][0]);
                    a[c][1]         = fma(c1, uc, a[c][1]);
                    a[c][2]         = fma(c2, uc, a[c][2]);
                  }
              }
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
              for (int b = 0; b < 3; ++b)
                F[(c * 3 + b) * m * n + gq * n + gi] = a[c][b];
          }
        __syncthreads();
        // integrate along y; accumulate the z integration of this level
        if (t_n2)
          {
#pragma unroll
            for (int c = 0; c < 3; ++c)
              {
                double        xy = 0, z = 0;
                const double *fc = F + (c * 3) * m * n + hi;
#pragma unroll L::UM
                for (int qy = 0; qy < m; ++qy)
                  {
                    const double s = sS[qy * n + hj];
                    const double g = sG[qy * n + hj];
                    xy = fma(s, fc[qy * n], xy);
                    xy = fma(g, fc[m * n + qy * n], xy);
                    z  = fma(s, fc[2 * m * n + qy * n], z);
                  }
#pragma unroll
                for (int k = 0; k < n; ++k)
                  acc[c][k] = fma(sS[qz * n + k], xy,
                                  fma(sG[qz * n + k], z, acc[c][k]));
              }
          }
      }

    // ---- inverse mass with QGauss(k+1): S^-1 diag(1/JxW) S^-T, lines of
    // the z-line owners (U is free); even-odd sweeps, S^-1 being
    // centrosymmetric ---------------------------------------------------------
    const DevBasis<double> &B = *Bd;
    // S^-T along z: (S^-T v)_q = sum_k Sinv[k][q] v_k
    if (r < n2)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        {
          double v[n];
          eo_apply<n, false, true, false>(B.eo_Sinv, acc[c], v);
#pragma unroll
          for (int q = 0; q < n; ++q)
            U[c * n3 + q * n2 + r] = v[q];
        }
    __syncthreads();
    // S^-T along y: thread (i, qz)
    if (r < n2)
      {
        const int i = r % n, qz = r / n;
#pragma unroll
        for (int c = 0; c < 3; ++c)
          {
            double v[n], o[n];
#pragma unroll
            for (int j = 0; j < n; ++j)
              v[j] = U[c * n3 + qz * n2 + j * n + i];
            eo_apply<n, false, true, false>(B.eo_Sinv, v, o);
#pragma unroll
            for (int q = 0; q < n; ++q)
              U[c * n3 + qz * n2 + q * n + i] = o[q];
          }
      }
    __syncthreads();
    // S^-T along x, 1/JxW, S^-1 along x: thread (qy, qz)
    if (r < n2)
      {
        const int qy = r % n, qz = r / n;
        double    ij[n];
        if constexpr (cartesian)
          {
            const double iw =
              B.winv[qy] * B.winv[qz] / (P.h[0] * P.h[1] * P.h[2]);
#pragma unroll
            for (int qx = 0; qx < n; ++qx)
              ij[qx] = B.winv[qx] * iw;
          }
        else
          {
#pragma unroll
            for (int qx = 0; qx < n; ++qx)
              {
                const double xi[3] = {B.xq[qx], B.xq[qy], B.xq[qz]};
                double       x[3], J[3][3], adj[3][3];
                multilinear_map<3>(X, xi, x, J);
                const double det = adjugate<3>(J, adj);
                ij[qx] = B.winv[qx] * B.winv[qy] * B.winv[qz] / det;
              }
          }
#pragma unroll
        for (int c = 0; c < 3; ++c)
          {
            double v[n], g[n], o[n];
#pragma unroll
            for (int i = 0; i < n; ++i)
              v[i] = U[c * n3 + qz * n2 + qy * n + i];
            eo_apply<n, false, true, false>(B.eo_Sinv, v, g);
#pragma unroll
            for (int q = 0; q < n; ++q)
              g[q] *= ij[q];
            eo_apply<n, false, false, false>(B.eo_Sinv, g, o);
#pragma unroll
            for (int i = 0; i < n; ++i)
              U[c * n3 + qz * n2 + qy * n + i] = o[i];
          }
      }
    __syncthreads();
    // S^-1 along y: thread (i, qz)
    if (r < n2)
      {
        const int i = r % n, qz = r / n;
#pragma unroll
        for (int c = 0; c < 3; ++c)
          {
            double v[n], o[n];
#pragma unroll
            for (int q = 0; q < n; ++q)
              v[q] = U[c * n3 + qz * n2 + q * n + i];
            eo_apply<n, false, false, false>(B.eo_Sinv, v, o);
#pragma unroll
            for (int j = 0; j < n; ++j)
              U[c * n3 + qz * n2 + j * n + i] = o[j];
          }
      }
    __syncthreads();
    // S^-1 along z and the epilogue: thread (i, j) owns its z-line
    if (r < n2 && active)
      {
#pragma unroll
        for (int c = 0; c < 3; ++c)
          {
            double v[n], o[n];
#pragma unroll
            for (int q = 0; q < n; ++q)
              v[q] = U[c * n3 + q * n2 + r];
            eo_apply<n, false, false, false>(B.eo_Sinv, v, o);
#pragma unroll
            for (int k = 0; k < n; ++k)
              {
                const int64_t g = cell * nvec + c * n3 + k * n2 + r;
                if constexpr (epi == epi_store)
                  io.dst[g] = io.scale * o[k];
                else
                  {
                    const double s0 = io.sol_in[g];
                    io.sol_out[g]   = fma(io.b_dt, o[k], s0);
                    if (!io.last)
                      io.ri_out[g] = fma(io.a_dt, o[k], s0);
                  }
              }
          }
      }
  }
} // namespace dgops
