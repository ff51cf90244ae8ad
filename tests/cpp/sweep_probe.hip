// Throughput of the 1D sweep at degree k (n = k + 1 points per line,
// FP64): the even-odd VALU products of the operator kernels (eo_apply, one
// line per thread) against FP64 matrix cores (v_mfma_f64_16x16x4f64, LPB =
// 16 / n lines block-diagonal per matrix column, 16 LPB lines per
// wavefront: 3 at degree 4, 2 at degree 7, where two 8 x 8 blocks fill the
// 16 rows). Both apply S and S^-1 alternately `reps` times to every line
// of a vector of 32,768,000 values (the size of the 64^3, degree-4 apply);
// reps = 0 is the memory time alone. Between two MFMA sweeps the result is
// turned back into the B-operand layout through a wavefront-private LDS
// tile, as a sweep in another direction would need.
//
// The lane layout of the MFMA operands is not assumed: a first kernel
// multiplies one-hot B operands with lane-numbered A operands, the host
// derives from the products which (row, k), (k, column) and (row, column)
// entries every lane holds and hands these tables to the sweep kernel.
// The results of both sweep kernels are compared on the host.
//
//     sweep_probe [-k degree] [reps ...]    degree 1..7, default 4;
//                                            reps default 0 8 32 128
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "dev_basis.hpp"
#include "op_kernel.hpp"

using namespace dgops;

namespace
{
  typedef double v4d __attribute__((ext_vector_type(4)));

  void
  check(hipError_t e, const char *what)
  {
    if (e != hipSuccess)
      {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
      }
  }

  // one wavefront: out[p][lane][i] = D of A (lane l holds l + 1) times B
  // (lane p holds 1, all others 0)
  __global__ void __launch_bounds__(64)
  mfma_layout(double *out)
  {
    const int l = threadIdx.x;
    for (int p = 0; p < 64; ++p)
      {
        const v4d d = __builtin_amdgcn_mfma_f64_16x16x4f64(
          (double)(l + 1), l == p ? 1. : 0., (v4d){0, 0, 0, 0}, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 4; ++i)
          out[(p * 64 + l) * 4 + i] = d[i];
      }
  }

  // per lane: A row, A k, B k, B column, D rows [4], D columns [4]
  struct Layout
  {
    int a_row[64], a_k[64], b_k[64], b_col[64], d_row[64][4], d_col[64][4];
  };

  // derives the layout from the products of mfma_layout; returns false if
  // they do not describe a 16 x 4 times 4 x 16 product
  bool
  derive_layout(const std::vector<double> &out, Layout &L)
  {
    std::map<std::set<int>, int> col_of_slots, k_of_alanes;
    std::vector<std::vector<std::pair<int, int>>> slots(64); // (slot, a lane)
    for (int p = 0; p < 64; ++p)
      {
        std::set<int> s, a;
        for (int l = 0; l < 64; ++l)
          for (int i = 0; i < 4; ++i)
            {
              const double v = out[(p * 64 + l) * 4 + i];
              if (v == 0.)
                continue;
              const int x = (int)v - 1;
              if (x < 0 || x > 63 || (double)(x + 1) != v)
                return false;
              s.insert(l * 4 + i);
              a.insert(x);
              slots[p].push_back({l * 4 + i, x});
            }
        if (s.size() != 16 || a.size() != 16)
          return false;
        if (!col_of_slots.count(s))
          {
            const int c = (int)col_of_slots.size();
            col_of_slots[s] = c;
          }
        if (!k_of_alanes.count(a))
          {
            const int k = (int)k_of_alanes.size();
            k_of_alanes[a] = k;
          }
        L.b_col[p] = col_of_slots[s];
        L.b_k[p]   = k_of_alanes[a];
      }
    if (col_of_slots.size() != 16 || k_of_alanes.size() != 4)
      return false;
    // rows: numbered by the A lanes of k = 0, in lane order
    std::vector<int> row_of_slot(256, -1), a_row(64, -1), a_k(64, -1);
    for (const auto &e : k_of_alanes)
      if (e.second == 0)
        {
          int r = 0;
          for (const int x : e.first)
            a_row[x] = r++;
        }
    for (int p = 0; p < 64; ++p)
      if (L.b_k[p] == 0)
        for (const auto &sx : slots[p])
          row_of_slot[sx.first] = a_row[sx.second];
    for (int p = 0; p < 64; ++p)
      for (const auto &sx : slots[p])
        {
          const int r = row_of_slot[sx.first];
          if (r < 0 || (a_row[sx.second] >= 0 && a_row[sx.second] != r))
            return false;
          a_row[sx.second] = r;
          a_k[sx.second]   = L.b_k[p];
        }
    std::set<std::pair<int, int>> a_seen, d_seen;
    for (int l = 0; l < 64; ++l)
      {
        if (a_row[l] < 0 || a_k[l] < 0)
          return false;
        L.a_row[l] = a_row[l];
        L.a_k[l]   = a_k[l];
        a_seen.insert({a_row[l], a_k[l]});
      }
    for (int p = 0; p < 64; ++p)
      for (const auto &sx : slots[p])
        {
          L.d_row[sx.first / 4][sx.first % 4] = row_of_slot[sx.first];
          L.d_col[sx.first / 4][sx.first % 4] = L.b_col[p];
          d_seen.insert({row_of_slot[sx.first], L.b_col[p]});
        }
    return a_seen.size() == 64 && d_seen.size() == 256;
  }

  template <int n>
  __global__ void __launch_bounds__(256)
  valu_sweeps(const DevBasis<double> *__restrict__ B, const double *in,
              double *out, int64_t lines, int reps)
  {
    const int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (l >= lines)
      return;
    double v[n], w[n];
#pragma unroll
    for (int i = 0; i < n; ++i)
      v[i] = in[l * n + i];
    for (int r = 0; r < reps; ++r)
      {
        if (r % 2 == 0)
          eo_apply<n, false, false, false>(B->eo_S, v, w);
        else
          eo_apply<n, false, false, false>(B->eo_Sinv, v, w);
#pragma unroll
        for (int i = 0; i < n; ++i)
          v[i] = w[i];
      }
#pragma unroll
    for (int i = 0; i < n; ++i)
      out[l * n + i] = v[i];
  }

  // Abd: [2][16][16] block-diagonal, LPB copies of S and of S^-1, zero in
  // the rows and columns from LPB n on; K-slice s of the product holds
  // columns 4 s .. 4 s + 3 of A. Column c of the B operand carries lines
  // LPB c .. LPB c + LPB - 1 of the wavefront's 16 LPB (row r = n b + j is
  // point j of line LPB c + b).
  template <int n>
  __global__ void __launch_bounds__(256)
  mfma_sweeps(const double *__restrict__ Abd, const Layout *__restrict__ L,
              const double *in, double *out, int64_t lines, int reps)
  {
    constexpr int LPB = 16 / n;
    __shared__ double tile[4][16 * 16];
    const int     lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    const int64_t l0   = ((int64_t)blockIdx.x * 4 + wv) * 16 * LPB;
    double       *T    = tile[wv];
    const int     ar = L->a_row[lane], ak = L->a_k[lane];
    const int     bk = L->b_k[lane], bc = L->b_col[lane];
    int           dr[4], dc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      {
        dr[i] = L->d_row[lane][i];
        dc[i] = L->d_col[lane][i];
      }
    double a[2][4], b[4];
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
      for (int s = 0; s < 4; ++s)
        a[m][s] = Abd[m * 256 + ar * 16 + 4 * s + ak];
#pragma unroll
    for (int s = 0; s < 4; ++s)
      {
        const int     r    = 4 * s + bk;
        const int64_t line = l0 + LPB * bc + r / n;
        b[s] = (r < LPB * n && line < lines) ? in[line * n + r % n] : 0.;
      }
    v4d d = {b[0], b[1], b[2], b[3]};
    for (int rep = 0; rep < reps; ++rep)
      {
        const int m = rep % 2;
        d           = (v4d){0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; ++s)
          d = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m][s], b[s], d, 0, 0,
                                                   0);
        // D layout -> B layout through the wavefront's own tile (the
        // wavefront's LDS accesses complete in order; the wave barriers
        // keep the compiler from moving them across each other)
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int i = 0; i < 4; ++i)
          T[dr[i] * 16 + dc[i]] = d[i];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int s = 0; s < 4; ++s)
          b[s] = T[(4 * s + bk) * 16 + bc];
        __builtin_amdgcn_wave_barrier();
      }
#pragma unroll
    for (int s = 0; s < 4; ++s)
      {
        const int     r    = 4 * s + bk;
        const int64_t line = l0 + LPB * bc + r / n;
        if (r < LPB * n && line < lines)
          out[line * n + r % n] = b[s];
      }
  }
} // namespace

namespace
{
  // times both sweep forms for n-point lines and compares their results;
  // returns 0, or 2 if they differ
  template <int n>
  int
  run(const Layout &L, const std::vector<int> &reps_list)
  {
    constexpr int LPB   = 16 / n;
    const int64_t lines = 32768000 / n;
    const Basis1D basis = make_basis(n);
    const auto    hb    = make_dev_basis<double>(basis);
    std::vector<double> Abd(2 * 256, 0.);
    for (int b = 0; b < LPB; ++b)
      for (int q = 0; q < n; ++q)
        for (int i = 0; i < n; ++i)
          {
            Abd[(n * b + q) * 16 + n * b + i]       = hb.S[q][i];
            Abd[256 + (n * b + q) * 16 + n * b + i] = hb.Sinv[q][i];
          }
    std::vector<double> h(lines * n);
    for (int64_t i = 0; i < lines * n; ++i)
      h[i] = std::sin(0.001 * (double)i);
    DevBasis<double> *dB;
    Layout           *dL;
    double           *dA, *din, *dv, *dm;
    check(hipMalloc(&dB, sizeof(hb)), "hipMalloc");
    check(hipMalloc(&dL, sizeof(L)), "hipMalloc");
    check(hipMalloc(&dA, sizeof(double) * Abd.size()), "hipMalloc");
    check(hipMalloc(&din, sizeof(double) * h.size()), "hipMalloc");
    check(hipMalloc(&dv, sizeof(double) * h.size()), "hipMalloc");
    check(hipMalloc(&dm, sizeof(double) * h.size()), "hipMalloc");
    check(hipMemcpy(dB, &hb, sizeof(hb), hipMemcpyHostToDevice), "copy");
    check(hipMemcpy(dL, &L, sizeof(L), hipMemcpyHostToDevice), "copy");
    check(hipMemcpy(dA, Abd.data(), sizeof(double) * Abd.size(),
                    hipMemcpyHostToDevice),
          "copy");
    check(hipMemcpy(din, h.data(), sizeof(double) * h.size(),
                    hipMemcpyHostToDevice),
          "copy");
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    const unsigned gv = (unsigned)((lines + 255) / 256);
    const unsigned gm = (unsigned)((lines + 64 * LPB - 1) / (64 * LPB));
    int            rc = 0;
    std::printf("degree %d: %d points per line, %d lines per MFMA block "
                "(%d of 16 rows), %lld lines\n",
                n - 1, n, LPB, LPB * n, (long long)lines);
    for (const int reps : reps_list)
      {
        float tv = 0, tm = 0;
        for (int it = 0; it < 4; ++it)
          {
            float t;
            check(hipEventRecord(e0, 0), "event");
            hipLaunchKernelGGL((valu_sweeps<n>), dim3(gv), dim3(256), 0, 0,
                               dB, din, dv, lines, reps);
            check(hipEventRecord(e1, 0), "event");
            check(hipEventSynchronize(e1), "sync");
            check(hipEventElapsedTime(&t, e0, e1), "time");
            if (it > 0)
              tv += t / 3;
            check(hipEventRecord(e0, 0), "event");
            hipLaunchKernelGGL((mfma_sweeps<n>), dim3(gm), dim3(256), 0, 0,
                               dA, dL, din, dm, lines, reps);
            check(hipEventRecord(e1, 0), "event");
            check(hipEventSynchronize(e1), "sync");
            check(hipEventElapsedTime(&t, e0, e1), "time");
            if (it > 0)
              tm += t / 3;
          }
        check(hipGetLastError(), "launch");
        std::vector<double> rv(h.size()), rm(h.size());
        check(hipMemcpy(rv.data(), dv, sizeof(double) * h.size(),
                        hipMemcpyDeviceToHost),
              "copy");
        check(hipMemcpy(rm.data(), dm, sizeof(double) * h.size(),
                        hipMemcpyDeviceToHost),
              "copy");
        double diff = 0, mx = 0;
        for (size_t i = 0; i < h.size(); ++i)
          {
            diff = std::fmax(diff, std::fabs(rv[i] - rm[i]));
            mx   = std::fmax(mx, std::fabs(rv[i]));
          }
        // line sweeps per second, and the dense-equivalent rate (2 n^2
        // flop per line and sweep)
        const double sweeps = (double)lines * reps;
        std::printf("reps %3d: VALU even-odd %.3f ms (%.2f G line-sweeps/s, "
                    "%.1f TFLOP/s dense-eq.), MFMA f64 16x16x4 %.3f ms "
                    "(%.2f, %.1f), max difference %.2e of %.2e\n",
                    reps, tv, sweeps / (tv * 1e-3) / 1e9,
                    2. * n * n * sweeps / (tv * 1e-3) / 1e12, tm,
                    sweeps / (tm * 1e-3) / 1e9,
                    2. * n * n * sweeps / (tm * 1e-3) / 1e12, diff, mx);
        if (!(diff <= 1e-11 * mx))
          rc = 2;
      }
    check(hipFree(dB), "hipFree");
    check(hipFree(dL), "hipFree");
    check(hipFree(dA), "hipFree");
    check(hipFree(din), "hipFree");
    check(hipFree(dv), "hipFree");
    check(hipFree(dm), "hipFree");
    return rc;
  }
} // namespace

int
main(int argc, char **argv)
{
  std::vector<int> reps_list;
  int              degree = 4;
  for (int i = 1; i < argc; ++i)
    {
      if (std::string(argv[i]) == "-k" && i + 1 < argc)
        degree = std::atoi(argv[++i]);
      else
        reps_list.push_back(std::atoi(argv[i]));
    }
  if (reps_list.empty())
    reps_list = {0, 8, 32, 128};
  if (degree < 1 || degree > 7)
    {
      std::fprintf(stderr, "degree must be 1..7\n");
      return 1;
    }

  double *dlay;
  check(hipMalloc(&dlay, sizeof(double) * 64 * 64 * 4), "hipMalloc");
  hipLaunchKernelGGL(mfma_layout, dim3(1), dim3(64), 0, 0, dlay);
  check(hipGetLastError(), "launch");
  std::vector<double> lay(64 * 64 * 4);
  check(hipMemcpy(lay.data(), dlay, sizeof(double) * lay.size(),
                  hipMemcpyDeviceToHost),
        "copy");
  Layout L;
  if (!derive_layout(lay, L))
    {
      std::fprintf(stderr, "the layout products are not a 16x4x16 product\n");
      return 3;
    }
  std::printf("v_mfma_f64_16x16x4f64 lanes 0, 1, 16, 17: A (row,k) "
              "(%d,%d) (%d,%d) (%d,%d) (%d,%d); B (k,col) (%d,%d) (%d,%d) "
              "(%d,%d) (%d,%d); D[0..3] (row,col) of lane 0 (%d,%d) (%d,%d) "
              "(%d,%d) (%d,%d), of lane 16 (%d,%d) ...\n",
              L.a_row[0], L.a_k[0], L.a_row[1], L.a_k[1], L.a_row[16],
              L.a_k[16], L.a_row[17], L.a_k[17], L.b_k[0], L.b_col[0],
              L.b_k[1], L.b_col[1], L.b_k[16], L.b_col[16], L.b_k[17],
              L.b_col[17], L.d_row[0][0], L.d_col[0][0], L.d_row[0][1],
              L.d_col[0][1], L.d_row[0][2], L.d_col[0][2], L.d_row[0][3],
              L.d_col[0][3], L.d_row[16][0], L.d_col[16][0]);
  switch (degree)
    {
      case 1: return run<2>(L, reps_list);
      case 2: return run<3>(L, reps_list);
      case 3: return run<4>(L, reps_list);
      case 4: return run<5>(L, reps_list);
      case 5: return run<6>(L, reps_list);
      case 6: return run<7>(L, reps_list);
      default: return run<8>(L, reps_list);
    }
}
