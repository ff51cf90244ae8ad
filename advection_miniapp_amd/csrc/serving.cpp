// Serving loop: many independent explicit advection solves resident on one
// GPU, admitted first come, first served within a device-memory budget and
// advanced together one time step per dg_server_step. All active solves of
// one degree share one kernel launch per LSRK stage (the batched operator
// kernel finds its solve from the workgroup index), so small solves fill
// the GPU together and a solve's arithmetic is the one it gets alone; the
// launches of different degrees run on streams of their own and overlap.
// Solves join and leave at step boundaries (continuous batching).
#include <algorithm>
#include <cmath>
#include <deque>
#include <iterator>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "aux.hpp"
#include "internal.hpp"
#include "registry.hpp"

namespace dgops
{
  std::map<uint32_t, BatchLaunch> &
  batch_registry()
  {
    static std::map<uint32_t, BatchLaunch> m;
    return m;
  }

  namespace
  {
    template <typename F>
    int
    guard(F &&f)
    {
      try
        {
          f();
          return DG_OK;
        }
      catch (const Error &e)
        {
          set_last_error(e.what());
          return e.code;
        }
      catch (const std::exception &e)
        {
          set_last_error(e.what());
          return DG_ERR_ARGUMENT;
        }
    }

    void
    require(bool c, const std::string &m, int code = DG_ERR_ARGUMENT)
    {
      if (!c)
        throw Error(code, m);
    }

    void
    request_config(const dg_serve_request &r, dg_config &cfg, double &dt,
                   int &n_steps)
    {
      require(r.degree >= 1 && r.degree <= 7, "degree must be 1..7",
              DG_ERR_UNSUPPORTED);
      int c = r.cells_per_direction, refine = 0;
      require(c >= 1, "cells_per_direction must be positive");
      while (c % 2 == 0)
        {
          c /= 2;
          ++refine;
        }
      require(c == 1 || c == 3 || c == 5 || c == 7,
              "cells_per_direction must be c 2^r with c in {1, 3, 5, 7}",
              DG_ERR_UNSUPPORTED);
      const double speed = std::sqrt(r.velocity[0] * r.velocity[0] +
                                     r.velocity[1] * r.velocity[1] +
                                     r.velocity[2] * r.velocity[2]);
      require(speed > 0 && r.final_time > 0 && r.courant > 0 && r.width > 0,
              "speed, final_time, courant and width must be positive");
      dg_config_default(&cfg);
      cfg.dim    = 3;
      cfg.degree = r.degree;
      for (int d = 0; d < 3; ++d)
        {
          cfg.n_coarse[d] = c;
          cfg.lo[d]       = 0.;
          cfg.hi[d]       = 1.;
          cfg.periodic[d] = 1;
          cfg.velocity[d] = r.velocity[d];
          cfg.solution_param[d] = r.center[d];
        }
      cfg.refine            = refine;
      cfg.velocity_kind     = DG_VEL_CONSTANT;
      cfg.solution_kind     = DG_SOL_GAUSSIAN;
      cfg.solution_param[3] = r.width;
      cfg.flux_alpha        = 1.;
      cfg.factor_skew       = 0.;
      // advection_solver.cc:1163-1167 with h_min = 1 / cells_per_direction
      const double h   = 1. / r.cells_per_direction;
      const double dt0 = r.courant * h / speed /
                         std::pow((double)std::max(1, r.degree), 1.5);
      n_steps = (int)std::ceil(r.final_time / dt0);
      dt      = r.final_time / n_steps;
    }

    struct ServeReq
    {
      int64_t          id;
      dg_serve_request req;
      dg_config        cfg;
      double           dt;
      int              n_steps, step = 0;
      int              state  = DG_REQUEST_QUEUED;
      dg_operator     *handle = nullptr;
      DevBuf           reg[3];
      int64_t          n_cells, n_dofs;
      int64_t          op_bytes, vec_bytes;
      int              err_code = DG_OK; // setup error of a FAILED request
      std::string      err_msg;
      int              shape = 0; // tile shape of its LSRK stage kernel

      Operator *op() const { return reinterpret_cast<Operator *>(handle); }
      ~ServeReq()
      {
        if (handle)
          dg_destroy(handle);
      }
    };
  } // namespace

  struct Server
  {
    hipStream_t stream = nullptr;
    int64_t     budget = 0, used = 0, next_id = 0;
    std::deque<int64_t>                          queue;
    std::vector<int64_t>                         active;
    std::map<int64_t, std::unique_ptr<ServeReq>> reqs;
    std::vector<LsrkStage>                       plan;
    DevBuf                                       d_items;
    BatchItem<double>                           *h_items = nullptr;
    size_t                                       h_cap   = 0;

    // one stream per degree group (at most max_lanes): the launches of
    // different degrees overlap on the device; each degree's stages stay in
    // order on its stream. Default priority, joined to `stream` by events.
    static constexpr int    max_lanes = 4;
    std::vector<hipStream_t> lanes;
    std::vector<hipEvent_t>  lane_done;
    hipEvent_t               ready = nullptr;

    ~Server()
    {
      if (h_items)
        (void)hipHostFree(h_items);
      for (hipStream_t l : lanes)
        (void)hipStreamDestroy(l);
      for (hipEvent_t e : lane_done)
        (void)hipEventDestroy(e);
      if (ready)
        (void)hipEventDestroy(ready);
    }

    void
    make_lanes(int n)
    {
      if (!ready)
        hip_check(hipEventCreateWithFlags(&ready, hipEventDisableTiming),
                  "serving event");
      while ((int)lanes.size() < n)
        {
          hipStream_t l;
          hipEvent_t  e;
          hip_check(hipStreamCreateWithFlags(&l, hipStreamNonBlocking),
                    "serving stream");
          lanes.push_back(l);
          hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming),
                    "serving event");
          lane_done.push_back(e);
        }
    }

    // device memory of a request: three registers, mesh arrays, bases
    static void
    footprint(ServeReq &q)
    {
      const int     n  = q.cfg.degree + 1;
      const int64_t N  = q.req.cells_per_direction;
      q.n_cells        = N * N * N;
      q.n_dofs         = q.n_cells * n * n * n;
      q.vec_bytes      = 8 * q.n_dofs;
      q.op_bytes       = q.n_cells * (6 * 4 + 8 * 3 * 8) +
                   (int64_t)(sizeof(DevBasis<double>) +
                             sizeof(DevBasis<float>));
    }

    // leave no half-built request behind: the caller sees the error of
    // the admitting call, the request's FAILED state, and the same error
    // again when it takes (and so releases) the request
    static void
    fail(ServeReq &q, int code, const std::string &msg)
    {
      for (auto &r : q.reg)
        r.release();
      if (q.handle)
        dg_destroy(q.handle);
      q.handle   = nullptr;
      q.state    = DG_REQUEST_FAILED;
      q.err_code = code;
      q.err_msg  = "request " + std::to_string(q.id) + " setup: " + msg;
    }

    // the tile shape of the request's fused LSRK stage kernel, tuned (or
    // found in the tuner's table) with its projected solution as input and
    // scratch outputs, as a solve run alone would tune it
    static int
    tuned_shape(ServeReq &q)
    {
      const OpParams    P  = q.op()->base_params(0., 1);
      OpBuffers<double> io = {};
      io.src     = q.reg[0].as<double>();
      io.sol_in  = q.reg[0].as<double>();
      io.sol_out = q.reg[1].as<double>();
      io.ri_out  = q.reg[2].as<double>();
      const uint32_t base = op_key(3, q.cfg.degree + 1, 1, 0, geo_uniform,
                                   0, 0, fin_inverse_mass, epi_lsrk);
      return tile_choice(*q.op(), base, P, false, epi_lsrk, &io).shape;
    }

    void
    admit()
    {
      while (!queue.empty())
        {
          ServeReq &q = *reqs.at(queue.front());
          const int64_t need = 3 * q.vec_bytes + q.op_bytes;
          if (used + need > budget)
            break;
          queue.pop_front();
          try
            {
              dg_operator *h  = nullptr;
              const int    rc = dg_create(&q.cfg, stream, &h);
              if (rc != DG_OK)
                throw Error(rc, dg_last_error());
              q.handle = h;
              for (auto &r : q.reg)
                r.alloc(q.vec_bytes);
              aux_project(*q.op(), q.reg[0].ptr, 0.);
              q.shape = tuned_shape(q);
            }
          catch (const Error &e)
            {
              fail(q, e.code, e.what());
              throw;
            }
          catch (const std::exception &e)
            {
              fail(q, DG_ERR_ARGUMENT, e.what());
              throw;
            }
          q.state = DG_REQUEST_ACTIVE;
          used += need;
          active.push_back(q.id);
        }
    }

    // advance all active solves by up to max_steps steps (stopping at the
    // first completion); returns the number of solves that finished
    int
    step(int max_steps)
    {
      admit();
      if (active.empty())
        return 0;
      int n_run = std::max(1, max_steps);
      for (int64_t id : active)
        {
          const ServeReq &q = *reqs.at(id);
          n_run = std::min(n_run, q.n_steps - q.step);
        }
      // groups of one degree, in admission order inside a group
      std::map<int, std::vector<ServeReq *>> groups;
      for (int64_t id : active)
        {
          ServeReq &q = *reqs.at(id);
          groups[q.cfg.degree + 1].push_back(&q);
        }
      const size_t n_items = (size_t)n_run * plan.size() * active.size();
      if (n_items > h_cap)
        {
          if (h_items)
            (void)hipHostFree(h_items);
          h_items = nullptr;
          hip_check(hipHostMalloc((void **)&h_items,
                                  sizeof(BatchItem<double>) * n_items),
                    "pinned descriptors");
          h_cap = n_items;
        }
      d_items.alloc(sizeof(BatchItem<double>) * n_items);
      struct Launch
      {
        BatchLaunch f;
        size_t      first;
        int         count;
        unsigned    blocks;
        const void *basis;
        int         lane;
      };
      const int n_lanes = std::min<int>(max_lanes, (int)groups.size());
      make_lanes(n_lanes);
      std::vector<Launch> launches;
      size_t              k = 0;
      for (int j = 0; j < n_run; ++j)
        for (const LsrkStage &s : plan)
          for (auto &g : groups)
            {
              const int lane = (int)std::distance(groups.begin(),
                                                  groups.find(g.first)) %
                               n_lanes;
              const int n = g.first;
              // the tile shape tuned for the group's largest solve
              const ServeReq *big = g.second.front();
              for (const ServeReq *q : g.second)
                if (q->n_cells > big->n_cells)
                  big = q;
              auto it = batch_registry().find(batch_key(3, n, big->shape));
              require(it != batch_registry().end(),
                      "batched operator not built for degree " +
                        std::to_string(n - 1),
                      DG_ERR_UNSUPPORTED);
              const int64_t cpb = tile_cells(3, n, big->shape);
              Launch        L   = {it->second, k, 0, 0,
                                   g.second.front()->op()->d_basis_d.ptr,
                                   lane};
              unsigned      first = 0;
              for (ServeReq *q : g.second)
                {
                  const double       t = (q->step + j) * q->dt;
                  BatchItem<double> &b = h_items[k++];
                  b                    = {};
                  b.P           = q->op()->base_params(t + s.c * q->dt, 1);
                  b.io.src      = q->reg[s.r_in].as<double>();
                  b.io.sol_in   = q->reg[s.s_in].as<double>();
                  b.io.sol_out  = q->reg[s.s_out].as<double>();
                  b.io.ri_out   = s.last ? nullptr : q->reg[s.r_out].as<double>();
                  b.io.b_dt     = s.b * q->dt;
                  b.io.a_dt     = s.a * q->dt;
                  b.io.last     = s.last;
                  b.first_block = first;
                  first += (unsigned)((q->n_cells + cpb - 1) / cpb);
                  ++L.count;
                }
              L.blocks = first;
              launches.push_back(L);
            }
      hip_check(hipMemcpyAsync(d_items.ptr, h_items,
                               sizeof(BatchItem<double>) * k,
                               hipMemcpyHostToDevice, stream),
                "upload descriptors");
      const auto *dev = d_items.as<BatchItem<double>>();
      // the lanes start after the descriptors (and earlier admissions)
      hip_check(hipEventRecord(ready, stream), "serving event");
      for (int l = 0; l < n_lanes; ++l)
        hip_check(hipStreamWaitEvent(lanes[l], ready, 0), "serving wait");
      for (const Launch &L : launches)
        hip_check(L.f(dev + L.first, L.count, L.blocks, L.basis,
                      lanes[L.lane]),
                  "batched operator launch");
      for (int l = 0; l < n_lanes; ++l)
        {
          hip_check(hipEventRecord(lane_done[l], lanes[l]), "serving event");
          hip_check(hipStreamWaitEvent(stream, lane_done[l], 0),
                    "serving wait");
        }
      // the pinned descriptors are reused by the next call
      hip_check(hipStreamSynchronize(stream), "serving step");
      // retire finished solves: keep the solution register only
      std::vector<int64_t> still;
      int                  finished = 0;
      for (int64_t id : active)
        {
          ServeReq &q = *reqs.at(id);
          q.step += n_run;
          if (q.step < q.n_steps)
            {
              still.push_back(id);
              continue;
            }
          q.state = DG_REQUEST_DONE;
          q.reg[1].release();
          q.reg[2].release();
          dg_destroy(q.handle);
          q.handle = nullptr;
          used -= 2 * q.vec_bytes + q.op_bytes;
          ++finished;
        }
      active.swap(still);
      return finished;
    }
  };
} // namespace dgops

using namespace dgops;

extern "C" {

int
dg_serve_request_config(const dg_serve_request *r, dg_config *cfg,
                        double *time_step, int *n_steps)
{
  return guard([&] {
    require(r && cfg && time_step && n_steps, "null argument");
    request_config(*r, *cfg, *time_step, *n_steps);
  });
}

int
dg_server_create(int64_t hbm_budget_bytes, void *stream, dg_server **out)
{
  return guard([&] {
    require(out != nullptr, "null output");
    require(hbm_budget_bytes > 0, "budget must be positive");
    auto s    = std::make_unique<Server>();
    s->stream = (hipStream_t)stream;
    s->budget = hbm_budget_bytes;
    s->plan   = lsrk_plan(DG_LSRK_5_4);
    *out      = reinterpret_cast<dg_server *>(s.release());
  });
}

int
dg_server_destroy(dg_server *s)
{
  return guard([&] { delete reinterpret_cast<Server *>(s); });
}

int
dg_server_submit(dg_server *hs, const dg_serve_request *r, int64_t *id)
{
  return guard([&] {
    require(hs && r && id, "null argument");
    Server &S = *reinterpret_cast<Server *>(hs);
    auto    q = std::make_unique<ServeReq>();
    q->req    = *r;
    request_config(*r, q->cfg, q->dt, q->n_steps);
    Server::footprint(*q);
    require(3 * q->vec_bytes + q->op_bytes <= S.budget,
            "request does not fit the memory budget");
    q->id = S.next_id++;
    *id   = q->id;
    S.queue.push_back(q->id);
    S.reqs[q->id] = std::move(q);
  });
}

int
dg_server_step(dg_server *hs, int max_steps, int *n_active, int *n_queued,
               int *n_finished)
{
  return guard([&] {
    require(hs != nullptr, "null server");
    Server   &S = *reinterpret_cast<Server *>(hs);
    const int f = S.step(max_steps);
    if (n_finished)
      *n_finished = f;
    if (n_active)
      *n_active = (int)S.active.size();
    if (n_queued)
      *n_queued = (int)S.queue.size();
  });
}

int
dg_server_request_info(dg_server *hs, int64_t id, int *state,
                       int64_t *n_dofs, int *steps_done, int *steps_total)
{
  return guard([&] {
    require(hs != nullptr, "null server");
    Server &S  = *reinterpret_cast<Server *>(hs);
    auto    it = S.reqs.find(id);
    if (it == S.reqs.end())
      {
        if (state)
          *state = DG_REQUEST_UNKNOWN;
        return;
      }
    const ServeReq &q = *it->second;
    if (state)
      *state = q.state;
    if (n_dofs)
      *n_dofs = q.n_dofs;
    if (steps_done)
      *steps_done = q.step;
    if (steps_total)
      *steps_total = q.n_steps;
  });
}

int
dg_server_take_result(dg_server *hs, int64_t id, void *dst)
{
  return guard([&] {
    require(hs && dst, "null argument");
    Server &S  = *reinterpret_cast<Server *>(hs);
    auto    it = S.reqs.find(id);
    require(it != S.reqs.end(), "unknown request");
    ServeReq &q = *it->second;
    if (q.state == DG_REQUEST_FAILED)
      {
        // release it and report its setup error
        const Error e(q.err_code, q.err_msg);
        S.reqs.erase(it);
        throw e;
      }
    require(q.state == DG_REQUEST_DONE, "request not finished",
            DG_ERR_STATE);
    hip_check(hipMemcpyAsync(dst, q.reg[0].ptr, q.vec_bytes,
                             hipMemcpyDeviceToDevice, S.stream),
              "copy result");
    hip_check(hipStreamSynchronize(S.stream), "copy result");
    S.used -= q.vec_bytes;
    S.reqs.erase(it);
  });
}

int
dg_server_memory(dg_server *hs, int64_t *bytes)
{
  return guard([&] {
    require(hs && bytes, "null argument");
    *bytes = reinterpret_cast<Server *>(hs)->used;
  });
}

} // extern "C"
