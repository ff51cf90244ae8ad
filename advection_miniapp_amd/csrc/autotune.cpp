// Per-degree tile autotuner of the operator kernels.
//
// A kernel variant (dimension, degree, components, precision, geometry
// mode, form, finalization, epilogue) is built in up to three tile shapes:
// one, two or four wavefronts of cell lines per workgroup. The first
// launch of a tunable variant on a device, for a mesh of a given size
// class, times every shape with HIP events on the operator's stream,
// writing into scratch vectors so that the caller's outputs are
// untouched, and keeps the fastest for the rest of the process. Every
// shape computes every cell with the same arithmetic, so the choice
// changes the time only, never the result. The reference leaves the
// batch shape to deal.II's MatrixFree (advection_solver.cc:436-465); this
// is the GPU's counterpart.
#include <map>
#include <mutex>
#include <tuple>

#include "internal.hpp"
#include "registry.hpp"

namespace dgops
{
  namespace
  {
    static_assert(n_tile_shapes == 3, "internal.hpp declares 3 shapes");
    // below this many cells the default shape is used untimed
    constexpr int64_t min_tuned_cells = 4096;
    // timed launches per candidate, after one warm-up launch
    constexpr int tune_reps = 5;

    struct Entry
    {
      TileChoice choice;
      double     ms[n_tile_shapes];
    };

    std::mutex &
    cache_mutex()
    {
      static std::mutex m;
      return m;
    }

    std::map<std::tuple<int, uint32_t, int>, Entry> &
    cache()
    {
      static std::map<std::tuple<int, uint32_t, int>, Entry> c;
      return c;
    }

    int
    size_class(int64_t n_cells)
    {
      int c = 0;
      while ((int64_t(1) << (c + 1)) <= n_cells)
        ++c;
      return c;
    }

    uint32_t
    with_shape(uint32_t base, int shape)
    {
      return base | (uint32_t)shape << 19;
    }

    // a copy of the launch's buffers whose outputs are scratch vectors
    template <typename T>
    OpBuffers<T>
    redirect(const void *io_in, int epi, T *a, T *b)
    {
      OpBuffers<T> io = *static_cast<const OpBuffers<T> *>(io_in);
      if (epi == epi_store)
        io.dst = a;
      else
        {
          io.sol_out = a;
          io.ri_out  = io.ri_out ? b : nullptr;
        }
      return io;
    }

    Entry
    tune(const Operator &op, uint32_t base, const OpParams &P, bool fp32,
         int epi, const void *io)
    {
      Entry        e  = {};
      const size_t es = fp32 ? 4 : 8;
      DevBuf       a, b;
      a.alloc(es * op.n_dofs);
      b.alloc(es * op.n_dofs);
      OpBuffers<double> iod = {};
      OpBuffers<float>  iof = {};
      if (fp32)
        iof = redirect<float>(io, epi, a.as<float>(), b.as<float>());
      else
        iod = redirect<double>(io, epi, a.as<double>(), b.as<double>());
      const void *io2 = fp32 ? (const void *)&iof : (const void *)&iod;
      hipEvent_t  t0, t1;
      hip_check(hipEventCreate(&t0), "tuner event");
      hip_check(hipEventCreate(&t1), "tuner event");
      double best = 1e300;
      for (int sh = 0; sh < n_tile_shapes; ++sh)
        {
          e.ms[sh] = -1;
          auto it  = op_registry().find(with_shape(base, sh));
          if (it == op_registry().end())
            continue;
          hip_check(it->second(P, op.basis_ptr(fp32), op.d_basis_d.ptr, io2,
                               op.stream),
                    "tuner warm-up launch");
          hip_check(hipEventRecord(t0, op.stream), "tuner event");
          for (int r = 0; r < tune_reps; ++r)
            hip_check(it->second(P, op.basis_ptr(fp32), op.d_basis_d.ptr,
                                 io2, op.stream),
                      "tuner launch");
          hip_check(hipEventRecord(t1, op.stream), "tuner event");
          hip_check(hipEventSynchronize(t1), "tuner sync");
          float ms = 0;
          hip_check(hipEventElapsedTime(&ms, t0, t1), "tuner time");
          e.ms[sh] = ms / tune_reps;
          if (e.ms[sh] < best)
            {
              best     = e.ms[sh];
              e.choice = {sh, e.ms[sh], true};
            }
        }
      (void)hipEventDestroy(t0);
      (void)hipEventDestroy(t1);
      return e;
    }
  } // namespace

  int
  tile_cells(int dim, int n, int shape)
  {
    const int nl = dim == 3 ? n * n : n;
    return shape_cells(nl, shape);
  }

  TileChoice
  tile_choice(Operator &op, uint32_t base, const OpParams &P, bool fp32,
              int epi, const void *io)
  {
    if (op.tile_force_cells > 0)
      {
        for (int sh = 0; sh < n_tile_shapes; ++sh)
          if (tile_cells(op.cfg.dim, op.n, sh) == op.tile_force_cells &&
              op_registry().count(with_shape(base, sh)))
            return {sh, 0., false};
        throw Error(DG_ERR_UNSUPPORTED,
                    "tile shape of " + std::to_string(op.tile_force_cells) +
                      " cells per workgroup not built for this operator "
                      "variant");
      }
    const bool tunable = op_registry().count(with_shape(base, 1)) &&
                         op_registry().count(with_shape(base, 2));
    if (!tunable || op.mesh.n_cells < min_tuned_cells)
      return {0, 0., false};
    const auto key =
      std::make_tuple(op.device, base, size_class(op.mesh.n_cells));
    std::lock_guard<std::mutex> lock(cache_mutex());
    auto                        it = cache().find(key);
    if (it == cache().end())
      it = cache().emplace(key, tune(op, base, P, fp32, epi, io)).first;
    return it->second.choice;
  }

  bool
  tile_candidates(const Operator &op, uint32_t base,
                  double (&ms)[n_tile_shapes])
  {
    std::lock_guard<std::mutex> lock(cache_mutex());
    auto it = cache().find(
      std::make_tuple(op.device, base, size_class(op.mesh.n_cells)));
    if (it == cache().end())
      return false;
    for (int s = 0; s < n_tile_shapes; ++s)
      ms[s] = it->second.ms[s];
    return true;
  }
} // namespace dgops

using namespace dgops;

extern "C" {

int
dg_set_tile_shape(dg_operator *h, int cells_per_workgroup)
{
  try
    {
      if (!h)
        throw Error(DG_ERR_ARGUMENT, "null operator");
      Operator &op = *reinterpret_cast<Operator *>(h);
      if (cells_per_workgroup < 0)
        throw Error(DG_ERR_ARGUMENT, "cells per workgroup must be >= 0");
      if (cells_per_workgroup > 0)
        {
          bool ok = false;
          for (int sh = 0; sh < n_tile_shapes; ++sh)
            ok = ok || tile_cells(op.cfg.dim, op.n, sh) == cells_per_workgroup;
          if (!ok)
            throw Error(DG_ERR_UNSUPPORTED,
                        "cells per workgroup must be one of the tile shapes "
                        "of this degree");
        }
      op.tile_force_cells = cells_per_workgroup;
      return DG_OK;
    }
  catch (const Error &e)
    {
      set_last_error(e.what());
      return e.code;
    }
}

int
dg_get_tile_shape(const dg_operator *h, int *cells_per_workgroup,
                  double *tuned_ms, double *candidates)
{
  try
    {
      if (!h)
        throw Error(DG_ERR_ARGUMENT, "null operator");
      const Operator &op = *reinterpret_cast<const Operator *>(h);
      if (op.last_tile_key == 0)
        throw Error(DG_ERR_STATE, "no operator launch on this handle yet");
      if (cells_per_workgroup)
        *cells_per_workgroup =
          tile_cells(op.cfg.dim, op.n, op.last_tile.shape);
      if (tuned_ms)
        *tuned_ms = op.last_tile.tuned ? op.last_tile.ms : 0.;
      if (candidates)
        {
          double     ms[n_tile_shapes];
          const bool t = tile_candidates(op, op.last_tile_key, ms);
          for (int sh = 0; sh < n_tile_shapes; ++sh)
            {
              candidates[2 * sh]     = tile_cells(op.cfg.dim, op.n, sh);
              candidates[2 * sh + 1] = t ? ms[sh] : -1.;
            }
        }
      return DG_OK;
    }
  catch (const Error &e)
    {
      set_last_error(e.what());
      return e.code;
    }
}

} // extern "C"
