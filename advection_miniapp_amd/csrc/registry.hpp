// Table of compiled kernel variants. Every generated translation unit
// registers its launchers at load time; the host code looks them up by key
// and reports DG_ERR_UNSUPPORTED for combinations that were not built.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>

#include "burgers3d_kernel.hpp"
#include "burgers_kernel.hpp"
#include "colloc_kernel.hpp"
#include "fdm_kernel.hpp"
#include "mass_kernel.hpp"
#include "op_kernel.hpp"

namespace dgops
{
  // launcher of one operator variant and tile shape
  using OpLaunch = hipError_t (*)(const OpParams &,
                                  const void *basis_t,
                                  const void *basis_d,
                                  const void *io,
                                  hipStream_t);

  inline uint32_t
  op_key(int dim, int n, int nc, int fp32, int geo, int skew, int diff,
         int fin, int epi, int shape = 0)
  {
    return (uint32_t)dim | (uint32_t)n << 2 | (uint32_t)nc << 6 |
           (uint32_t)fp32 << 10 | (uint32_t)geo << 11 |
           (uint32_t)skew << 13 | (uint32_t)fin << 14 | (uint32_t)epi << 16 |
           (uint32_t)diff << 18 | (uint32_t)shape << 19;
  }

  std::map<uint32_t, OpLaunch> &op_registry();

  struct OpRegistrar
  {
    OpRegistrar(uint32_t key, OpLaunch f) { op_registry()[key] = f; }
  };

  // FDM preconditioner variants: prec 0 = fp64, 1 = fp32, 2 = fp32
  // arithmetic on fp64 vectors
  using FdmLaunch = hipError_t (*)(const FdmParams &,
                                   const void *tables,
                                   const void *src,
                                   void       *dst,
                                   hipStream_t);

  inline uint32_t
  fdm_key(int dim, int n, int ns, int prec)
  {
    return (uint32_t)dim | (uint32_t)n << 2 | (uint32_t)ns << 6 |
           (uint32_t)prec << 10 | 1u << 20;
  }

  std::map<uint32_t, FdmLaunch> &fdm_registry();

  struct FdmRegistrar
  {
    FdmRegistrar(uint32_t key, FdmLaunch f) { fdm_registry()[key] = f; }
  };

  template <int dim, int n, int NS, typename Tc, typename Tin, typename Tout>
  hipError_t
  launch_fdm(const FdmParams &P,
             const void      *tables,
             const void      *src,
             void            *dst,
             hipStream_t      stream)
  {
    constexpr int CPB    = cells_per_block<dim, n>();
    constexpr int nl     = Lay<dim, n>::nl;
    static_assert(CPB * nl <= 128, "kernels are bounded to 128 threads");
    const int64_t blocks = (P.n_cells + CPB - 1) / CPB;
    if (blocks <= 0)
      return hipSuccess;
    hipLaunchKernelGGL((fdm_kernel<dim, n, NS, Tc, Tin, Tout>),
                       dim3((unsigned)blocks),
                       dim3(CPB * nl),
                       0,
                       stream,
                       P,
                       (const FdmDev<Tc> *)tables,
                       (const Tin *)src,
                       (Tout *)dst);
    return hipGetLastError();
  }

  // Burgers right-hand side variants
  using BurgersLaunch = hipError_t (*)(const OpParams &,
                                       const void *basis_d,
                                       const void *tables,
                                       const void *io,
                                       hipStream_t);

  inline uint32_t
  burgers_key(int dim, int n, int cart, int epi)
  {
    return (uint32_t)dim | (uint32_t)n << 2 | (uint32_t)cart << 6 |
           (uint32_t)epi << 7 | 1u << 21;
  }

  std::map<uint32_t, BurgersLaunch> &burgers_registry();

  struct BurgersRegistrar
  {
    BurgersRegistrar(uint32_t key, BurgersLaunch f)
    {
      burgers_registry()[key] = f;
    }
  };

  template <int dim, int n, int m, bool cart, int epi>
  hipError_t
  launch_burgers(const OpParams &P,
                 const void     *bd,
                 const void     *tables,
                 const void     *io,
                 hipStream_t     stream)
  {
    if (P.n_cells <= 0)
      return hipSuccess;
    hipLaunchKernelGGL((burgers_kernel<dim, n, m, double, cart, epi>),
                       dim3((unsigned)P.n_cells),
                       dim3(256),
                       0,
                       stream,
                       P,
                       (const DevBasis<double> *)bd,
                       (const BurgersTables *)tables,
                       *(const OpBuffers<double> *)io);
    return hipGetLastError();
  }

  // 3D Burgers, degrees 1..7 (slab-wise cell term)
  template <int dim, int n, int m, bool cart, int epi>
  hipError_t
  launch_burgers3(const OpParams &P,
                  const void     *bd,
                  const void     *tables,
                  const void     *io,
                  hipStream_t     stream)
  {
    static_assert(dim == 3, "the slab-wise Burgers kernel is 3D");
    constexpr int CPB = burgers3_cells(n, m);
    constexpr int R   = burgers3_threads(n, m);
    static_assert(CPB * R <= 1024, "workgroup size");
    const int64_t blocks = (P.n_cells + CPB - 1) / CPB;
    if (blocks <= 0)
      return hipSuccess;
    hipLaunchKernelGGL((burgers3_kernel<n, m, cart, epi, CPB>),
                       dim3((unsigned)blocks), dim3(CPB * R), 0, stream, P,
                       (const DevBasis<double> *)bd,
                       (const BurgersTables *)tables,
                       *(const OpBuffers<double> *)io);
    return hipGetLastError();
  }

  // cellwise inverse mass matrix variants (in place)
  using MassLaunch = hipError_t (*)(const OpParams &, const void *basis_t,
                                    const void *basis_d, void *vec,
                                    hipStream_t);

  inline uint32_t
  mass_key(int dim, int n, int fp32)
  {
    return (uint32_t)dim | (uint32_t)n << 2 | (uint32_t)fp32 << 6 | 1u << 22;
  }

  std::map<uint32_t, MassLaunch> &mass_registry();

  struct MassRegistrar
  {
    MassRegistrar(uint32_t key, MassLaunch f) { mass_registry()[key] = f; }
  };

  template <int dim, int n, typename T>
  hipError_t
  launch_mass(const OpParams &P, const void *bt, const void *bd, void *vec,
              hipStream_t stream)
  {
    constexpr int CPB    = cells_per_block<dim, n>();
    constexpr int nl     = Lay<dim, n>::nl;
    static_assert(CPB * nl <= 128, "kernels are bounded to 128 threads");
    const int64_t blocks = (P.n_cells + CPB - 1) / CPB;
    if (blocks <= 0)
      return hipSuccess;
    hipLaunchKernelGGL((inverse_mass_line_kernel<dim, n, T>),
                       dim3((unsigned)blocks), dim3(CPB * nl), 0, stream, P,
                       (const DevBasis<T> *)bt, (const DevBasis<double> *)bd,
                       (T *)vec);
    return hipGetLastError();
  }

  // batched cell solver of the IRK preconditioner: nodal -> collocation
  // values, and the per-cell GMRES update with the transform back
  using ToCollocLaunch = hipError_t (*)(int64_t n_cells, int64_t block,
                                        int ns, const void *basis,
                                        const void *src, void *dst,
                                        hipStream_t);
  using CellGmresLaunch = hipError_t (*)(int64_t n_cells, int64_t block,
                                         const void *basis, const void *r,
                                         const void *y, const void *z,
                                         void *dst, hipStream_t);

  inline uint32_t
  colloc_key(int dim, int n, int fp32)
  {
    return (uint32_t)dim | (uint32_t)n << 2 | (uint32_t)fp32 << 6 | 1u << 23;
  }

  inline uint32_t
  cell_gmres_key(int dim, int n, int ns, int fp32)
  {
    return (uint32_t)dim | (uint32_t)n << 2 | (uint32_t)ns << 6 |
           (uint32_t)fp32 << 10 | 1u << 24;
  }

  std::map<uint32_t, ToCollocLaunch>  &colloc_registry();
  std::map<uint32_t, CellGmresLaunch> &cell_gmres_registry();

  struct CollocRegistrar
  {
    CollocRegistrar(uint32_t key, ToCollocLaunch f)
    {
      colloc_registry()[key] = f;
    }
  };

  struct CellGmresRegistrar
  {
    CellGmresRegistrar(uint32_t key, CellGmresLaunch f)
    {
      cell_gmres_registry()[key] = f;
    }
  };

  template <int dim, int n, typename T>
  hipError_t
  launch_to_colloc(int64_t n_cells, int64_t block, int ns, const void *bt,
                   const void *src, void *dst, hipStream_t stream)
  {
    constexpr int CPB = cells_per_block<dim, n>();
    constexpr int nl  = Lay<dim, n>::nl;
    static_assert(CPB * nl <= 128, "kernels are bounded to 128 threads");
    const int64_t blocks = (n_cells + CPB - 1) / CPB;
    if (blocks <= 0 || ns <= 0)
      return hipSuccess;
    hipLaunchKernelGGL((to_collocation_kernel<dim, n, T>),
                       dim3((unsigned)blocks, (unsigned)ns), dim3(CPB * nl),
                       0, stream, n_cells, block, (const DevBasis<T> *)bt,
                       (const T *)src, (T *)dst);
    return hipGetLastError();
  }

  template <int dim, int n, int NS, typename T>
  hipError_t
  launch_cell_gmres(int64_t n_cells, int64_t block, const void *bt,
                    const void *r, const void *y, const void *z, void *dst,
                    hipStream_t stream)
  {
    constexpr int CPB = cells_per_block<dim, n>();
    constexpr int nl  = Lay<dim, n>::nl;
    static_assert(CPB * nl <= 128, "kernels are bounded to 128 threads");
    const int64_t blocks = (n_cells + CPB - 1) / CPB;
    if (blocks <= 0)
      return hipSuccess;
    hipLaunchKernelGGL((cell_gmres_update_kernel<dim, n, NS, T>),
                       dim3((unsigned)blocks), dim3(CPB * nl), 0, stream,
                       n_cells, block, (const DevBasis<T> *)bt,
                       (const T *)r, (const T *)y, (const T *)z, (T *)dst);
    return hipGetLastError();
  }

  // batched explicit LSRK stages of independent axis-aligned,
  // constant-speed operators of one degree (serving loop)
  using BatchLaunch = hipError_t (*)(const void *items, int n_items,
                                     unsigned blocks, const void *basis,
                                     hipStream_t stream);

  inline uint32_t
  batch_key(int dim, int n, int shape = 0)
  {
    return (uint32_t)dim | (uint32_t)n << 2 | (uint32_t)shape << 6 |
           1u << 21;
  }

  std::map<uint32_t, BatchLaunch> &batch_registry();

  struct BatchRegistrar
  {
    BatchRegistrar(uint32_t key, BatchLaunch f) { batch_registry()[key] = f; }
  };

  template <int dim, int n, int CPB>
  hipError_t
  launch_batch(const void *items, int n_items, unsigned blocks,
               const void *basis, hipStream_t stream)
  {
    constexpr int nl = Lay<dim, n>::nl;
    static_assert(CPB * nl <= 256, "kernels are bounded to 256 threads");
    if (blocks == 0 || n_items <= 0)
      return hipSuccess;
    hipLaunchKernelGGL((dg_operator_batch_kernel<dim, n, 1, double,
                                                 geo_uniform, false, false,
                                                 fin_inverse_mass, epi_lsrk,
                                                 CPB>),
                       dim3(blocks), dim3(CPB * nl), 0, stream,
                       (const BatchItem<double> *)items, n_items,
                       (const DevBasis<double> *)basis,
                       (const DevBasis<double> *)basis);
    return hipGetLastError();
  }

  template <int dim, int n, int NC, typename T, int geo, bool skew,
            bool diff, int fin, int epi, int CPB>
  hipError_t
  launch_op(const OpParams &P,
            const void     *bt,
            const void     *bd,
            const void     *io,
            hipStream_t     stream)
  {
    constexpr int nl     = Lay<dim, n>::nl;
    static_assert(CPB * nl <= 256, "kernels are bounded to 256 threads");
    const int64_t blocks = (P.n_cells + CPB - 1) / CPB;
    if (blocks <= 0)
      return hipSuccess;
    hipLaunchKernelGGL(
      (dg_operator_kernel<dim, n, NC, T, geo, skew, diff, fin, epi, CPB>),
      dim3((unsigned)blocks),
      dim3(CPB * nl),
      0,
      stream,
      P,
      (const DevBasis<T> *)bt,
      (const DevBasis<double> *)bd,
      *(const OpBuffers<T> *)io);
    return hipGetLastError();
  }
} // namespace dgops
