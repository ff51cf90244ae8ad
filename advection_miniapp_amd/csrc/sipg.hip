// Launcher of the SIPG diffusion kernel for cells of any shape
// (sipg_kernel.hpp), one instance per dimension and degree; the face
// geometry table is built on the first call of an operator handle.
#include "aux.hpp"
#include "internal.hpp"
#include "sipg_kernel.hpp"

namespace dgops
{
  namespace
  {
    template <int dim, int n>
    void
    launch(Operator &op, const OpParams &P, const DevBasis<double> *B,
           const double *src, double *dst, double scale, int64_t n_cells,
           hipStream_t stream)
    {
      using L = SipgLay<dim, n>;
      if (!op.sipg_faces_ready)
        {
          const int64_t entries = n_cells * dim * L::nf;
          op.d_sipg_faces.alloc(sizeof(double) * entries * L::nt);
          hipLaunchKernelGGL((sipg_face_table_kernel<dim, n>),
                             dim3((entries + 255) / 256), dim3(256), 0,
                             stream, P, B, op.d_sipg_faces.as<double>());
          hip_check(hipGetLastError(), "SIPG face table kernel");
          op.sipg_faces_ready = true;
        }
      const int64_t blocks = (n_cells + L::CPB - 1) / L::CPB;
      hipLaunchKernelGGL((sipg_general_kernel<dim, n>), dim3(blocks),
                         dim3(L::NT), 0, stream, P, B,
                         op.d_sipg_faces.as<const double>(), src, dst, scale);
    }

    template <int dim, int n>
    bool
    dispatch(int want, Operator &op, const OpParams &P,
             const DevBasis<double> *B, const double *src, double *dst,
             double scale, int64_t n_cells, hipStream_t stream)
    {
      if (want != n)
        return false;
      launch<dim, n>(op, P, B, src, dst, scale, n_cells, stream);
      return true;
    }

    template <int dim, int... ns>
    bool
    dispatch_all(int n, Operator &op, const OpParams &P,
                 const DevBasis<double> *B, const double *src, double *dst,
                 double scale, int64_t n_cells, hipStream_t stream)
    {
      return (dispatch<dim, ns>(n, op, P, B, src, dst, scale, n_cells,
                                stream) ||
              ...);
    }
  } // namespace

  void
  aux_sipg_general(Operator &op, const OpParams &P, const void *src,
                   void *dst, double scale)
  {
    if (op.mesh.n_cells == 0)
      return;
    const auto *B  = op.d_basis_d.as<DevBasis<double>>();
    const auto *s  = static_cast<const double *>(src);
    auto       *d  = static_cast<double *>(dst);
    const bool  ok =
      op.cfg.dim == 3 ?
        dispatch_all<3, 2, 3, 4, 5, 6, 7, 8>(op.n, op, P, B, s, d, scale,
                                             op.mesh.n_cells, op.stream) :
        dispatch_all<2, 2, 3, 4, 5, 6, 7, 8>(op.n, op, P, B, s, d, scale,
                                             op.mesh.n_cells, op.stream);
    if (!ok)
      throw Error(DG_ERR_UNSUPPORTED, "SIPG on general cells: degree 1..7");
    hip_check(hipGetLastError(), "SIPG diffusion kernel");
  }
} // namespace dgops
