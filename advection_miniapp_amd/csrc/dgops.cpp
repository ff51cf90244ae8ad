// C ABI of the library: handle management, the explicit operator and the
// low-storage Runge-Kutta integrator, index export, vector operations.
#include <cmath>
#include <cstring>
#include <map>

#include "aux.hpp"
#include "internal.hpp"
#include "registry.hpp"

namespace dgops
{
  std::map<uint32_t, OpLaunch> &
  op_registry()
  {
    static std::map<uint32_t, OpLaunch> m;
    return m;
  }

  std::map<uint32_t, ToCollocLaunch> &
  colloc_registry()
  {
    static std::map<uint32_t, ToCollocLaunch> m;
    return m;
  }

  std::map<uint32_t, CellGmresLaunch> &
  cell_gmres_registry()
  {
    static std::map<uint32_t, CellGmresLaunch> m;
    return m;
  }

  std::map<uint32_t, MassLaunch> &
  mass_registry()
  {
    static std::map<uint32_t, MassLaunch> m;
    return m;
  }

  std::map<uint32_t, BurgersLaunch> &
  burgers_registry()
  {
    static std::map<uint32_t, BurgersLaunch> m;
    return m;
  }

  // Burgers right-hand side (burgers_solver.cc:259-341) with the fused
  // LSRK epilogue or a plain store
  static void
  launch_burgers_op(const Operator &op, const OpParams &P, int epi,
                    const void *io)
  {
    auto it = burgers_registry().find(
      burgers_key(op.cfg.dim, op.n, P.cartesian ? 1 : 0, epi));
    if (it == burgers_registry().end())
      throw Error(DG_ERR_UNSUPPORTED,
                  "Burgers operator variant not built for this degree");
    hip_check(it->second(P, op.d_basis_d.ptr, op.d_burgers.ptr, io,
                         op.stream),
              "burgers launch");
  }

  namespace
  {
    thread_local std::string last_error;

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
          last_error = e.what();
          return e.code;
        }
      catch (const std::exception &e)
        {
          last_error = e.what();
          return DG_ERR_ARGUMENT;
        }
      catch (...)
        {
          last_error = "unknown error";
          return DG_ERR_ARGUMENT;
        }
    }

    void
    require(bool cond, const std::string &msg, int code = DG_ERR_ARGUMENT)
    {
      if (!cond)
        throw Error(code, msg);
    }

    Operator *
    get(dg_operator *op)
    {
      require(op != nullptr, "null operator handle");
      return reinterpret_cast<Operator *>(op);
    }
    const Operator *
    get(const dg_operator *op)
    {
      require(op != nullptr, "null operator handle");
      return reinterpret_cast<const Operator *>(op);
    }

    // Low-storage Runge-Kutta coefficients (b_i, a_i) of the schemes of
    // advection_solver.cc:828-897 (Kennedy, Carpenter, Lewis 2000;
    // Tselios, Simos 2007)
    void
    lsrk_coefficients(int scheme, std::vector<double> &bi,
                      std::vector<double> &ai)
    {
      switch (scheme)
        {
          case DG_LSRK_3_3:
            bi = {0.245170287303492, 0.184896052186740, 0.569933660509768};
            ai = {0.755726351946097, 0.386954477304099};
            break;
          case DG_LSRK_5_4:
            bi = {1153189308089. / 22510343858157.,
                  1772645290293. / 4653164025191.,
                  -1672844663538. / 4480602732383.,
                  2114624349019. / 3568978502595.,
                  5198255086312. / 14908931495163.};
            ai = {970286171893. / 4311952581923.,
                  6584761158862. / 12103376702013.,
                  2251764453980. / 15575788980749.,
                  26877169314380. / 34165994151039.};
            break;
          case DG_LSRK_7_4:
            bi = {0.0941840925477795334, 0.149683694803496998,
                  0.285204742060440058,  -0.122201846148053668,
                  0.0605151571191401122, 0.345986987898399296,
                  0.186627171718797670};
            ai = {0.241566650129646868 + bi[0], 0.0423866513027719953 + bi[1],
                  0.215602732678803776 + bi[2], 0.232328007537583987 + bi[3],
                  0.256223412574146438 + bi[4], 0.0978694102142697230 + bi[5]};
            break;
          case DG_LSRK_9_5:
            bi = {2274579626619. / 23610510767302.,
                  693987741272. / 12394497460941.,
                  -347131529483. / 15096185902911.,
                  1144057200723. / 32081666971178.,
                  1562491064753. / 11797114684756.,
                  13113619727965. / 44346030145118.,
                  393957816125. / 7825732611452.,
                  720647959663. / 6565743875477.,
                  3559252274877. / 14424734981077.};
            ai = {1107026461565. / 5417078080134.,
                  38141181049399. / 41724347789894.,
                  493273079041. / 11940823631197.,
                  1851571280403. / 6147804934346.,
                  11782306865191. / 62590030070788.,
                  9452544825720. / 13648368537481.,
                  4435885630781. / 26285702406235.,
                  2357909744247. / 11371140753790.};
            break;
          default:
            throw Error(DG_ERR_ARGUMENT, "unknown LSRK scheme");
        }
    }
  } // namespace

  void
  set_last_error(const std::string &msg)
  {
    last_error = msg;
  }

  double
  velocity_time_factor(const dg_config &cfg, double t)
  {
    if (cfg.time_factor == DG_TIME_FACTOR_COS)
      return std::cos(pi * t / cfg.final_time);
    return 1.;
  }

  // Three registers rotate so that the stage input is never written by the
  // fused kernel that reads it; the last stage writes register 0 (the
  // solution). Stage times c follow advection_solver.cc:924-937.
  std::vector<LsrkStage>
  lsrk_plan(int scheme)
  {
    std::vector<double> bi, ai;
    lsrk_coefficients(scheme, bi, ai);
    const int              ns = (int)bi.size();
    std::vector<LsrkStage> plan;
    int                    r_in = 0, s_in = 0;
    double                 sum_previous_bi = 0;
    for (int stage = 0; stage < ns; ++stage)
      {
        LsrkStage s = {};
        s.c         = stage == 0 ? 0. : sum_previous_bi + ai[stage - 1];
        if (stage > 0)
          sum_previous_bi += bi[stage - 1];
        s.b    = bi[stage];
        s.last = stage == ns - 1;
        s.r_in = r_in;
        s.s_in = s_in;
        if (s.last)
          {
            s.a     = 0.;
            s.s_out = 0;
            s.r_out = -1;
            plan.push_back(s);
            break;
          }
        s.a = ai[stage];
        int cand[2], k = 0;
        for (int b = 0; b < 3; ++b)
          if (b != r_in)
            cand[k++] = b;
        if (stage == ns - 2)
          {
            // the next (last) stage reads r_out and writes register 0
            s.r_out = cand[0] == 0 ? cand[1] : cand[0];
            s.s_out = cand[0] == s.r_out ? cand[1] : cand[0];
          }
        else if (cand[0] == s_in || cand[1] == s_in)
          {
            s.s_out = s_in;
            s.r_out = cand[0] == s_in ? cand[1] : cand[0];
          }
        else
          {
            s.s_out = cand[0];
            s.r_out = cand[1];
          }
        plan.push_back(s);
        r_in = s.r_out;
        s_in = s.s_out;
      }
    return plan;
  }

  OpParams
  Operator::base_params(double t, int n_comp) const
  {
    OpParams P = {};
    P.n_cells       = mesh.n_cells;
    P.neighbors     = d_neighbors.as<int>();
    P.cell_vertices = d_vertices.as<double>();
    for (int d = 0; d < 3; ++d)
      {
        P.h[d]         = mesh.h[d];
        P.vel_const[d] = cfg.velocity[d];
      }
    P.cartesian = cfg.perturbation == 0.;
    P.vel_kind  = cfg.velocity_kind;
    const double tf = velocity_time_factor(cfg, t);
    for (int c = 0; c < 8; ++c)
      P.vel_factor[c] = c < n_comp ? tf : 0.;
    P.alpha    = cfg.flux_alpha;
    P.skew     = cfg.factor_skew;
    P.sign     = -1.;
    P.has_mass = 0;
    P.sol_kind = cfg.solution_kind;
    P.sol_time = t;
    for (int i = 0; i < 8; ++i)
      P.sol_param[i] = cfg.solution_param[i];
    P.viscosity = cfg.viscosity;
    P.penalty   = cfg.penalty_factor;
    // one periodic 3D coarse cell: neighbours from the Morton code
    const bool one_periodic_cube =
      cfg.dim == 3 && cfg.refine <= 10 && cfg.n_coarse[0] == 1 &&
      cfg.n_coarse[1] == 1 && cfg.n_coarse[2] == 1 && cfg.periodic[0] &&
      cfg.periodic[1] && cfg.periodic[2];
    P.morton_bits = one_periodic_cube ? cfg.refine : -1;
    return P;
  }

  void *
  Operator::solver_vector(size_t idx, size_t bytes)
  {
    while (solver_vectors.size() <= idx)
      solver_vectors.emplace_back(new DevBuf());
    solver_vectors[idx]->alloc(bytes);
    return solver_vectors[idx]->ptr;
  }

  void
  launch_operator(const Operator &op,
                  const OpParams &P,
                  int             nc,
                  bool            fp32,
                  bool            skew,
                  int             fin,
                  int             epi,
                  const void     *io,
                  bool            advection_only)
  {
    const int diff =
      op.cfg.operator_kind == DG_OP_ADVECTION_DIFFUSION && !advection_only;
    if (diff && !P.cartesian)
      throw Error(DG_ERR_UNSUPPORTED,
                  "the fused SIPG operator variants are built for "
                  "axis-aligned meshes (perturbation 0)");
    // axis-aligned cells with a constant transport speed need no tables
    // (products of 1D weights); SIPG with a variable speed evaluates it per
    // point; everything else reads the precomputed geometry tables
    const bool uniform =
      P.cartesian && op.cfg.velocity_kind == DG_VEL_CONSTANT;
    const int geo = uniform ? geo_uniform : diff ? geo_analytic : geo_table;
    OpParams  Pk  = P;
    Operator &m   = const_cast<Operator &>(op);
    if (geo == geo_table)
      {
        aux_geometry_tables(m, fp32 ? DG_FP32 : DG_FP64);
        Pk.tab_cell = m.d_tab_cell[fp32 ? 1 : 0].ptr;
        Pk.tab_face = m.d_tab_face[fp32 ? 1 : 0].ptr;
      }
    else
      aux_boundary_tables(m);
    Pk.tab_bnd   = m.d_tab_bnd.as<double>();
    Pk.bnd_index = m.d_bnd_index.as<int>();
    Pk.bnd_value = m.d_bnd_value.as<double>();
    aux_boundary_values(m, Pk);
    const uint32_t base = op_key(op.cfg.dim, op.n, nc, fp32 ? 1 : 0,
                                 geo, skew ? 1 : 0, diff, fin, epi);
    const TileChoice tile = tile_choice(m, base, Pk, fp32, epi, io);
    m.last_tile           = tile;
    m.last_tile_key       = base;
    const uint32_t key    = base | (uint32_t)tile.shape << 19;
    auto it = op_registry().find(key);
    if (it == op_registry().end())
      throw Error(DG_ERR_UNSUPPORTED,
                  "operator variant not built: dim=" +
                    std::to_string(op.cfg.dim) +
                    " degree=" + std::to_string(op.n - 1) +
                    " components=" + std::to_string(nc) +
                    (fp32 ? " fp32" : " fp64") +
                    (geo == geo_uniform    ? " uniform geometry"
                     : geo == geo_analytic ? " analytic geometry"
                                           : " tabulated geometry") +
                    (skew ? " skew" : "") + " fin=" + std::to_string(fin) +
                    " epi=" + std::to_string(epi));
    hip_check(it->second(Pk, op.basis_ptr(fp32), op.d_basis_d.ptr, io,
                         op.stream),
              "operator launch");
  }

  namespace
  {
    MeshDesc
    mesh_desc(const dg_config &cfg)
    {
      MeshDesc md = {};
      md.dim      = cfg.dim;
      for (int d = 0; d < 3; ++d)
        {
          md.n_coarse[d] = d < cfg.dim ? cfg.n_coarse[d] : 1;
          md.lo[d]       = cfg.lo[d];
          md.hi[d]       = cfg.hi[d];
          md.periodic[d] = cfg.periodic[d];
        }
      md.refine       = cfg.refine;
      md.perturbation = cfg.perturbation;
      md.seed         = cfg.seed;
      return md;
    }

    void
    validate(const dg_config &c)
    {
      require(c.dim == 2 || c.dim == 3, "dim must be 2 or 3");
      require(c.degree >= 1 && c.degree <= max_n - 1, "degree must be 1..7",
              DG_ERR_UNSUPPORTED);
      require(c.precision == DG_FP64 || c.precision == DG_FP32,
              "bad precision");
      require(c.operator_kind >= 0 && c.operator_kind <= 2,
              "bad operator kind");
      if (c.operator_kind == DG_OP_BURGERS)
        require(c.n_components == c.dim, "Burgers needs dim components");
      else
        require(c.n_components == 1, "scalar operator needs 1 component");
      require(c.refine >= 0 && c.refine <= 12, "refine out of range");
      for (int d = 0; d < c.dim; ++d)
        require(c.n_coarse[d] >= 1 && c.hi[d] > c.lo[d], "bad mesh box");
      // cell and neighbour indices are 32-bit
      double cells = std::ldexp(1.0, c.dim * c.refine);
      for (int d = 0; d < c.dim; ++d)
        cells *= c.n_coarse[d];
      require(cells <= 2147483647.,
              "mesh too large: at most 2^31 - 1 cells per operator",
              DG_ERR_UNSUPPORTED);
      require(c.perturbation >= 0 && c.perturbation < 0.5,
              "perturbation must be in [0, 0.5)");
      require(c.time_factor == DG_TIME_FACTOR_ONE || c.final_time > 0,
              "final_time must be positive for the cos time factor");
      require(c.irk_stages >= 0 && c.irk_stages <= 5, "irk_stages 0..5");
      require((c.use_gl_quad == 0 || c.use_gl_quad == 1) &&
                (c.use_gl_quad_mass == 0 || c.use_gl_quad_mass == 1),
              "use_gl_quad and use_gl_quad_mass are 0 or 1");
      if (c.operator_kind == DG_OP_BURGERS)
        require(!c.use_gl_quad && !c.use_gl_quad_mass,
                "Burgers operator: QGauss quadratures only (burgers_solver.cc "
                "has no quadrature switch)",
                DG_ERR_UNSUPPORTED);
      if (c.velocity_kind == DG_VEL_CELLULAR_3D)
        require(c.dim == 3, "cellular velocity field is 3D");
      if (c.operator_kind == DG_OP_ADVECTION_DIFFUSION)
        for (int d = 0; d < c.dim; ++d)
          require(c.periodic[d],
                  "advection-diffusion: periodic meshes only (the reference "
                  "has no boundary faces for SIPG)",
                  DG_ERR_UNSUPPORTED);
      if (c.operator_kind == DG_OP_BURGERS)
        for (int d = 0; d < c.dim; ++d)
          require(c.periodic[d],
                  "Burgers: periodic meshes only (the reference's boundary "
                  "face term is not implemented, burgers_solver.cc:347-355)",
                  DG_ERR_UNSUPPORTED);
    }
  } // namespace

  // the apply in separate passes (advection_solver.cc:661-691: the loop,
  // then the inverse mass): the operator's integration S^T (...) into dst,
  // the SIPG diffusion terms on perturbed cells added to it, then the
  // inverse mass matrix in place. Used when the mass matrix is integrated
  // by QGaussLobatto(k+1) and the operator by QGauss(k+1) (use_gl_quad_mass
  // alone; the inverse mass is then diag(1 / JxW) at the nodes, :443-445),
  // and for SIPG on perturbed cells.
  bool
  two_pass_apply(const Operator &op)
  {
    return op.gl_mass_only || op.sipg_general();
  }

  void
  apply_two_pass(Operator &op, const OpParams &P, const void *src, void *dst)
  {
    const bool general = op.sipg_general();
    if (op.fp32())
      {
        require(!general,
                "SIPG on perturbed cells is built in FP64 (the FP32 "
                "variants need axis-aligned cells)",
                DG_ERR_UNSUPPORTED);
        OpBuffers<float> io = {};
        io.src   = (const float *)src;
        io.dst   = (float *)dst;
        io.scale = 1.f;
        launch_operator(op, P, 1, true, true, fin_integrate, epi_store, &io);
      }
    else
      {
        OpBuffers<double> io = {};
        io.src   = (const double *)src;
        io.dst   = (double *)dst;
        io.scale = 1.;
        launch_operator(op, P, 1, false, true, fin_integrate, epi_store, &io,
                        general);
        if (general)
          aux_sipg_general(op, P, src, dst, 1.);
      }
    aux_inverse_mass(op, op.cfg.precision, dst);
  }
} // namespace dgops

using namespace dgops;

extern "C" {

const char *
dg_last_error(void)
{
  return last_error.c_str();
}

void
dg_config_default(dg_config *c)
{
  if (!c)
    return;
  std::memset(c, 0, sizeof(*c));
  c->dim           = 3;
  c->degree        = 4;
  c->n_components  = 1;
  c->operator_kind = DG_OP_ADVECTION;
  c->precision     = DG_FP64;
  for (int d = 0; d < 3; ++d)
    {
      c->n_coarse[d] = 1;
      c->lo[d]       = 0.;
      c->hi[d]       = 1.;
      c->periodic[d] = 1;
    }
  c->refine         = 2;
  c->perturbation   = 0.;
  c->seed           = 42;
  c->velocity_kind  = DG_VEL_CONSTANT;
  c->velocity[0]    = 1.;
  c->time_factor    = DG_TIME_FACTOR_ONE;
  c->final_time     = 1.;
  c->flux_alpha     = 1.;
  c->factor_skew    = 0.;
  c->viscosity      = 0.;
  c->penalty_factor = 0.;
  c->solution_kind  = DG_SOL_ZERO;
  c->irk_stages     = 3;
}

int
dg_create(const dg_config *cfg, void *stream, dg_operator **out)
{
  return guard([&] {
    require(cfg && out, "null argument");
    *out = nullptr;
    validate(*cfg);
    auto op    = std::make_unique<Operator>();
    op->cfg    = *cfg;
    op->stream = (hipStream_t)stream;
    hip_check(hipGetDevice(&op->device), "hipGetDevice");
    op->mesh        = make_mesh(mesh_desc(*cfg));
    op->n           = cfg->degree + 1;
    op->nd          = cfg->dim == 3 ? op->n * op->n * op->n : op->n * op->n;
    op->n_dofs      = op->mesh.n_cells * op->nd;
    op->basis       = make_basis(op->n, cfg->use_gl_quad != 0);
    op->gl_mass_only = !cfg->use_gl_quad && cfg->use_gl_quad_mass;
    if (cfg->operator_kind == DG_OP_BURGERS)
      {
        // QGauss(k(k+1)/2) tables (burgers_solver.cc:234)
        require(cfg->degree <= (cfg->dim == 3 ? 7 : 4),
                "Burgers operator: degree 1..7 in 3D, 1..4 in 2D",
                DG_ERR_UNSUPPORTED);
        require(cfg->precision == DG_FP64, "Burgers operator: FP64",
                DG_ERR_UNSUPPORTED);
        for (int d = 0; d < cfg->dim; ++d)
          require(cfg->periodic[d],
                  "Burgers operator: periodic meshes only (the reference "
                  "has no boundary faces)",
                  DG_ERR_UNSUPPORTED);
        BurgersTables T = {};
        const int     m = cfg->degree * (cfg->degree + 1) / 2;
        T.m             = m;
        std::vector<double> gp, gw;
        gauss_points(m, gp, gw);
        const auto xl = gauss_lobatto_points_ext(op->n);
        for (int q = 0; q < m; ++q)
          {
            T.wm[q] = gw[q];
            T.xm[q] = gp[q];
            for (int i = 0; i < op->n; ++i)
              {
                real_ext v, dv;
                lagrange(xl, i, (real_ext)gp[q], v, dv);
                T.Sm[q][i] = (double)v;
                T.Gm[q][i] = (double)dv;
              }
          }
        op->d_burgers.alloc(sizeof(T));
        hip_check(hipMemcpy(op->d_burgers.ptr, &T, sizeof(T),
                            hipMemcpyHostToDevice),
                  "upload burgers tables");
      }

    const auto bd = make_dev_basis<double>(op->basis);
    const auto bf = make_dev_basis<float>(op->basis);
    op->d_basis_d.alloc(sizeof(bd));
    op->d_basis_f.alloc(sizeof(bf));
    hip_check(hipMemcpy(op->d_basis_d.ptr, &bd, sizeof(bd),
                        hipMemcpyHostToDevice),
              "upload basis");
    hip_check(hipMemcpy(op->d_basis_f.ptr, &bf, sizeof(bf),
                        hipMemcpyHostToDevice),
              "upload basis");
    if (op->gl_mass_only)
      {
        const Basis1D mb  = make_basis(op->n, true);
        const auto    mbd = make_dev_basis<double>(mb);
        const auto    mbf = make_dev_basis<float>(mb);
        op->d_mbasis_d.alloc(sizeof(mbd));
        op->d_mbasis_f.alloc(sizeof(mbf));
        hip_check(hipMemcpy(op->d_mbasis_d.ptr, &mbd, sizeof(mbd),
                            hipMemcpyHostToDevice),
                  "upload mass basis");
        hip_check(hipMemcpy(op->d_mbasis_f.ptr, &mbf, sizeof(mbf),
                            hipMemcpyHostToDevice),
                  "upload mass basis");
      }
    const size_t nb = op->mesh.neighbors.size() * sizeof(int);
    const size_t nv = op->mesh.cell_vertices.size() * sizeof(double);
    op->d_neighbors.alloc(nb);
    op->d_vertices.alloc(nv);
    hip_check(hipMemcpy(op->d_neighbors.ptr, op->mesh.neighbors.data(), nb,
                        hipMemcpyHostToDevice),
              "upload neighbors");
    hip_check(hipMemcpy(op->d_vertices.ptr, op->mesh.cell_vertices.data(), nv,
                        hipMemcpyHostToDevice),
              "upload vertices");
    *out = reinterpret_cast<dg_operator *>(op.release());
  });
}

int
dg_destroy(dg_operator *op)
{
  return guard([&] { delete reinterpret_cast<Operator *>(op); });
}

int
dg_set_stream(dg_operator *h, void *stream)
{
  return guard([&] { get(h)->stream = (hipStream_t)stream; });
}

int
dg_get_info(const dg_operator *h, dg_info *info)
{
  return guard([&] {
    const Operator *op = get(h);
    require(info != nullptr, "null info");
    info->n_cells       = op->mesh.n_cells;
    info->n_dofs        = op->n_dofs;
    info->vector_size   = op->n_dofs * op->cfg.n_components;
    info->dofs_per_cell = op->nd;
    for (int d = 0; d < 3; ++d)
      {
        info->cells_per_direction[d] = op->mesh.N[d];
        info->h[d]                   = op->mesh.h[d];
      }
    info->min_vertex_distance = op->mesh.min_vertex_distance;
  });
}

int
dg_apply(dg_operator *h, const void *src, void *dst, double t)
{
  return guard([&] {
    Operator *op = get(h);
    require(src && dst, "null vector");
    require(src != dst, "dg_apply: src and dst must differ");
    if (op->cfg.operator_kind == DG_OP_BURGERS)
      {
        OpParams          P  = op->base_params(t, 1);
        OpBuffers<double> io = {};
        io.src   = (const double *)src;
        io.dst   = (double *)dst;
        io.scale = 1.;
        launch_burgers_op(*op, P, epi_store, &io);
        return;
      }
    // skew-symmetric form (factor_skew, advection_solver.cc:106-108,
    // :500-507): built for advection; SIPG uses the conservative form
    const bool skew = op->cfg.factor_skew != 0.;
    require(!skew || op->cfg.operator_kind == DG_OP_ADVECTION,
            "explicit advection-diffusion operator is built for "
            "factor_skew = 0",
            DG_ERR_UNSUPPORTED);
    OpParams P = op->base_params(t, 1);
    if (two_pass_apply(*op))
      {
        apply_two_pass(*op, P, src, dst);
        return;
      }
    if (op->fp32())
      {
        OpBuffers<float> io = {};
        io.src   = (const float *)src;
        io.dst   = (float *)dst;
        io.scale = 1.f;
        launch_operator(*op, P, 1, true, skew, fin_inverse_mass, epi_store,
                        &io);
      }
    else
      {
        OpBuffers<double> io = {};
        io.src   = (const double *)src;
        io.dst   = (double *)dst;
        io.scale = 1.;
        launch_operator(*op, P, 1, false, skew, fin_inverse_mass, epi_store,
                        &io);
      }
  });
}

namespace dgops
{
  // fused stage: reads `ri_in` (all of it, including neighbours) and
  // `sol_in` (own entries); writes sol_out and ri_out. ri_in must not alias
  // an output.
  static void
  fused_stage(Operator &op, double t, double b_dt, double a_dt,
              const void *ri_in, const void *sol_in, void *sol_out,
              void *ri_out, bool last)
  {
    OpParams P = op.base_params(t, 1);
    if (op.cfg.operator_kind == DG_OP_BURGERS)
      {
        OpBuffers<double> io = {};
        io.src     = (const double *)ri_in;
        io.sol_in  = (const double *)sol_in;
        io.sol_out = (double *)sol_out;
        io.ri_out  = (double *)ri_out;
        io.b_dt    = b_dt;
        io.a_dt    = a_dt;
        io.last    = last;
        launch_burgers_op(op, P, epi_lsrk, &io);
        return;
      }
    const bool skew = op.cfg.factor_skew != 0.;
    require(!skew || op.cfg.operator_kind == DG_OP_ADVECTION,
            "explicit advection-diffusion operator is built for "
            "factor_skew = 0",
            DG_ERR_UNSUPPORTED);
    // the LSRK stage accumulates in FP64; FP32 operators give right-hand
    // sides (apply) only
    require(!op.fp32(),
            "perform_stage: the fused LSRK stage is built in FP64 (create "
            "the operator with precision DG_FP64)",
            DG_ERR_UNSUPPORTED);
    if (two_pass_apply(op))
      {
        // two passes: k = M^-1 (-L ri_in), then the stage update
        const size_t bytes = sizeof(double) * op.n_dofs;
        op.d_gl_k.alloc(bytes);
        apply_two_pass(op, P, ri_in, op.d_gl_k.ptr);
        vec_lsrk_update(op, DG_FP64, op.n_dofs, b_dt, a_dt, op.d_gl_k.ptr,
                        sol_in, sol_out, ri_out);
        return;
      }
    {
        OpBuffers<double> io = {};
        io.src     = (const double *)ri_in;
        io.sol_in  = (const double *)sol_in;
        io.sol_out = (double *)sol_out;
        io.ri_out  = (double *)ri_out;
        io.b_dt    = b_dt;
        io.a_dt    = a_dt;
        io.last    = last;
        launch_operator(op, P, 1, false, skew, fin_inverse_mass, epi_lsrk,
                        &io);
      }
  }
} // namespace dgops

int
dg_perform_stage(dg_operator *h, double t, double b_dt, double a_dt,
                 const void *current_ri, void *vec_ki, void *solution,
                 void *next_ri)
{
  return guard([&] {
    Operator *op = get(h);
    require(current_ri && vec_ki && solution && next_ri, "null vector");
    const bool last = a_dt == 0.;
    if (current_ri != solution && current_ri != next_ri)
      {
        fused_stage(*op, t, b_dt, a_dt, current_ri, solution, solution,
                    last ? nullptr : next_ri, last);
        return;
      }
    // aliased registers as in LowStorageRungeKuttaIntegrator: two passes
    // through vec_ki like the reference (loop, then inverse mass + update)
    require(vec_ki != current_ri,
            "perform_stage: vec_ki must not alias current_ri");
    require(dg_apply(h, current_ri, vec_ki, t) == DG_OK, last_error);
    vec_lsrk_update(*op, op->cfg.precision,
                    op->n_dofs * op->cfg.n_components, b_dt, a_dt, vec_ki,
                    solution, solution, last ? nullptr : next_ri);
  });
}

int
dg_lsrk_n_stages(int scheme)
{
  std::vector<double> b, a;
  try
    {
      lsrk_coefficients(scheme, b, a);
    }
  catch (...)
    {
      return -1;
    }
  return (int)b.size();
}

int
dg_lsrk_coefficients(int scheme, double *bi, double *ai)
{
  std::vector<double> b, a;
  try
    {
      lsrk_coefficients(scheme, b, a);
    }
  catch (...)
    {
      return -1;
    }
  if (!bi || !ai)
    return -1;
  for (size_t i = 0; i < b.size(); ++i)
    {
      bi[i] = b[i];
      ai[i] = i < a.size() ? a[i] : 0.;
    }
  return (int)b.size();
}

namespace dgops
{
  // one LSRK step: the stages of the plan, fused (the register rotation
  // ends with the solution back in `solution`)
  static void
  lsrk_step(Operator &op, int scheme, double t, double dt, void *const (&buf)[3])
  {
    for (const LsrkStage &s : lsrk_plan(scheme))
      fused_stage(op, t + s.c * dt, s.b * dt, s.a * dt, buf[s.r_in],
                  buf[s.s_in], buf[s.s_out], s.last ? nullptr : buf[s.r_out],
                  s.last);
  }

  // the operator does not depend on time: constant speed factor and no
  // boundary faces (whose Dirichlet data follow the stage time)
  static bool
  autonomous(const Operator &op)
  {
    if (op.cfg.time_factor != DG_TIME_FACTOR_ONE)
      return false;
    for (const int nb : op.mesh.neighbors)
      if (nb < 0)
        return false;
    return true;
  }

  // the cached graph of DG_GRAPH_STEPS steps for (scheme, dt, registers),
  // captured on a stream of its own (the caller's stream may be the null
  // stream); the launches are the same as eager ones, so replays give the
  // same bits
  static hipGraphExec_t
  step_graph(Operator &op, int scheme, double dt, void *const (&buf)[3])
  {
    auto &g = op.step_graph;
    if (g.exec && g.scheme == scheme && g.dt == dt && g.reg[0] == buf[0] &&
        g.reg[1] == buf[1] && g.reg[2] == buf[2])
      return g.exec;
    if (g.exec)
      {
        (void)hipGraphExecDestroy(g.exec);
        g.exec = nullptr;
      }
    hipStream_t cap;
    hip_check(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking),
              "capture stream");
    hipStream_t saved = op.stream;
    hipGraph_t  graph = nullptr;
    op.stream         = cap;
    try
      {
        hip_check(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal),
                  "begin capture");
        try
          {
            for (int s = 0; s < DG_GRAPH_STEPS; ++s)
              lsrk_step(op, scheme, 0., dt, buf);
          }
        catch (...)
          {
            hipGraph_t g0 = nullptr;
            (void)hipStreamEndCapture(cap, &g0);
            if (g0)
              (void)hipGraphDestroy(g0);
            throw;
          }
        hip_check(hipStreamEndCapture(cap, &graph), "end capture");
        hip_check(hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0),
                  "graph instantiate");
      }
    catch (...)
      {
        op.stream = saved;
        if (graph)
          (void)hipGraphDestroy(graph);
        (void)hipStreamDestroy(cap);
        throw;
      }
    op.stream = saved;
    (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(cap);
    g.scheme = scheme;
    g.dt     = dt;
    for (int i = 0; i < 3; ++i)
      g.reg[i] = buf[i];
    return g.exec;
  }

  struct StepEvents
  {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    StepEvents()
    {
      for (auto &x : e)
        hip_check(hipEventCreate(&x), "step timing event");
    }
    ~StepEvents()
    {
      for (auto &x : e)
        if (x)
          (void)hipEventDestroy(x);
    }
  };

  // steps [s, n) of an autonomous operator: graph replays of
  // DG_GRAPH_STEPS steps where the launch form is (or is measured to be)
  // the graph, the rest one by one. The measurement spends
  // 3 DG_GRAPH_STEPS real steps: one replay to warm the graph, then
  // DG_GRAPH_STEPS steps in each form between events.
  static void
  autonomous_steps(Operator &op, int scheme, double t, double dt, int s,
                   int n, void *const (&buf)[3])
  {
    auto     &g    = op.step_graph;
    const int mode = g.force ? g.force : g.mode;
    if (mode == 0 && n - s >= 3 * DG_GRAPH_STEPS)
      {
        const hipGraphExec_t exec = step_graph(op, scheme, dt, buf);
        hip_check(hipGraphLaunch(exec, op.stream), "graph launch");
        s += DG_GRAPH_STEPS;
        StepEvents ev;
        hip_check(hipEventRecord(ev.e[0], op.stream), "step timing event");
        for (int i = 0; i < DG_GRAPH_STEPS; ++i, ++s)
          lsrk_step(op, scheme, t + s * dt, dt, buf);
        hip_check(hipEventRecord(ev.e[1], op.stream), "step timing event");
        hip_check(hipGraphLaunch(exec, op.stream), "graph launch");
        s += DG_GRAPH_STEPS;
        hip_check(hipEventRecord(ev.e[2], op.stream), "step timing event");
        hip_check(hipEventSynchronize(ev.e[2]), "step timing sync");
        hip_check(hipEventElapsedTime(&g.eager_ms, ev.e[0], ev.e[1]),
                  "step timing");
        hip_check(hipEventElapsedTime(&g.graph_ms, ev.e[1], ev.e[2]),
                  "step timing");
        g.mode = g.graph_ms < g.eager_ms ? 2 : 1;
      }
    if ((g.force ? g.force : g.mode) == 2)
      for (; n - s >= DG_GRAPH_STEPS; s += DG_GRAPH_STEPS)
        hip_check(hipGraphLaunch(step_graph(op, scheme, dt, buf), op.stream),
                  "graph launch");
    for (; s < n; ++s)
      lsrk_step(op, scheme, t + s * dt, dt, buf);
  }
} // namespace dgops

int
dg_lsrk_time_step(dg_operator *h, int scheme, double t, double dt,
                  void *solution, void *vec_ri, void *vec_ki)
{
  return guard([&] {
    Operator *op = get(h);
    require(solution && vec_ri && vec_ki, "null vector");
    require(solution != vec_ri && solution != vec_ki && vec_ri != vec_ki,
            "the three registers must be distinct");
    void *const buf[3] = {solution, vec_ri, vec_ki};
    lsrk_step(*op, scheme, t, dt, buf);
  });
}

int
dg_lsrk_time_steps(dg_operator *h, int scheme, double t, double dt,
                   int n_steps, void *solution, void *vec_ri, void *vec_ki)
{
  return guard([&] {
    Operator *op = get(h);
    require(solution && vec_ri && vec_ki, "null vector");
    require(solution != vec_ri && solution != vec_ki && vec_ri != vec_ki,
            "the three registers must be distinct");
    require(n_steps >= 0, "n_steps must be >= 0");
    void *const buf[3] = {solution, vec_ri, vec_ki};
    if (n_steps == 0)
      return;
    // the first step eagerly: it builds the tables, tunes the tile shape
    // and allocates the workspace, none of which a captured step may do
    lsrk_step(*op, scheme, t, dt, buf);
    if (n_steps == 1)
      return;
    if (!autonomous(*op))
      {
        for (int s = 1; s < n_steps; ++s)
          lsrk_step(*op, scheme, t + s * dt, dt, buf);
        return;
      }
    autonomous_steps(*op, scheme, t, dt, 1, n_steps, buf);
  });
}

int
dg_set_step_launch(dg_operator *h, int mode)
{
  return guard([&] {
    Operator *op = get(h);
    require(mode >= 0 && mode <= 2,
            "launch form must be 0 (measure), 1 (stream) or 2 (graph)");
    op->step_graph.force = mode;
  });
}

int
dg_get_step_launch(const dg_operator *h, int *mode, double *eager_ms,
                   double *graph_ms)
{
  return guard([&] {
    const Operator *op = get(h);
    const auto     &g  = op->step_graph;
    if (mode)
      *mode = g.force ? g.force : g.mode;
    if (eager_ms)
      *eager_ms = g.eager_ms;
    if (graph_ms)
      *graph_ms = g.graph_ms;
  });
}

int
dg_project_initial(dg_operator *h, void *dst, double t)
{
  return guard([&] {
    Operator *op = get(h);
    require(dst != nullptr, "null vector");
    aux_project(*op, dst, t);
  });
}

int
dg_compute_mass_and_energy(dg_operator *h, const void *vec, double out[3])
{
  return guard([&] {
    Operator *op = get(h);
    require(vec && out, "null argument");
    require(op->cfg.n_components == 1, "scalar fields only");
    aux_mass_energy(*op, vec, out);
  });
}

int
dg_l2_error(dg_operator *h, const void *vec, double t, double out[2])
{
  return guard([&] {
    Operator *op = get(h);
    require(vec && out, "null argument");
    require(op->cfg.n_components == 1, "scalar fields only");
    aux_l2_error(*op, vec, t, out);
  });
}

int
dg_vector_dot(dg_operator *h, int p, int64_t n, const void *x, const void *y,
              double *result)
{
  return guard([&] {
    require(x && y && result, "null argument");
    vec_multi_dot(*get(h), p, n, 1, x, n, y, result);
  });
}

int
dg_vector_multi_dot(dg_operator *h, int p, int64_t n, int m, const void *x,
                    const void *y, double *result)
{
  return guard([&] {
    require(x && y && result && m >= 0, "bad argument");
    vec_multi_dot(*get(h), p, n, m, x, n, y, result);
  });
}

int
dg_vector_multi_dot_noalloc(dg_operator *h, int p, int64_t n, int m,
                            const void *x, const void *y, double *result)
{
  return guard([&] {
    require(x && y && result && m >= 0, "bad argument");
    require(p == DG_FP64 || p == DG_FP32, "precision must be FP64 or FP32");
    vec_multi_dot_noalloc(*get(h), p, n, m, x, n, y, result);
  });
}

int
dg_vector_multi_axpy(dg_operator *h, int p, int64_t n, int m,
                     const double *alpha, const void *x, void *y)
{
  return guard([&] {
    require(x && y && alpha && m >= 0, "bad argument");
    vec_multi_axpy(*get(h), p, n, m, alpha, x, n, y);
  });
}

int
dg_vector_axpby(dg_operator *h, int p, int64_t n, double a, const void *x,
                double b, void *y)
{
  return guard([&] {
    require(x && y, "null vector");
    vec_axpby(*get(h), p, n, a, x, b, y);
  });
}

int
dg_vector_copy_convert(dg_operator *h, int64_t n, int ps, const void *src,
                       int pd, void *dst)
{
  return guard([&] {
    require(src && dst, "null vector");
    vec_copy_convert(*get(h), n, ps, src, pd, dst);
  });
}

int
dg_export_cell_coordinates(const dg_operator *h, int32_t *ijk)
{
  return guard([&] {
    const Operator *op = get(h);
    require(ijk != nullptr, "null output");
    for (size_t i = 0; i < op->mesh.cell_ijk.size(); ++i)
      ijk[i] = op->mesh.cell_ijk[i];
  });
}

int
dg_export_neighbors(const dg_operator *h, int32_t *nb)
{
  return guard([&] {
    const Operator *op = get(h);
    require(nb != nullptr, "null output");
    for (size_t i = 0; i < op->mesh.neighbors.size(); ++i)
      nb[i] = op->mesh.neighbors[i];
  });
}

int
dg_export_cell_vertices(const dg_operator *h, double *v)
{
  return guard([&] {
    const Operator *op = get(h);
    require(v != nullptr, "null output");
    std::memcpy(v, op->mesh.cell_vertices.data(),
                op->mesh.cell_vertices.size() * sizeof(double));
  });
}

int
dg_write_vtu(dg_operator *h, const void *vec, double t, const char *filename)
{
  return guard([&] {
    Operator *op = get(h);
    require(vec && filename, "null argument");
    const int64_t count = op->n_dofs * op->cfg.n_components;
    std::vector<double> host(count);
    hip_check(hipStreamSynchronize(op->stream), "sync");
    if (op->fp32())
      {
        std::vector<float> f(count);
        hip_check(hipMemcpy(f.data(), vec, sizeof(float) * count,
                            hipMemcpyDeviceToHost),
                  "copy solution");
        for (int64_t i = 0; i < count; ++i)
          host[i] = f[i];
      }
    else
      hip_check(hipMemcpy(host.data(), vec, sizeof(double) * count,
                          hipMemcpyDeviceToHost),
                "copy solution");
    write_vtu_host(op->cfg, op->mesh, host.data(), t, filename);
  });
}

int
dg_host_write_vtu(const dg_config *cfg, const double *vec, double t,
                  const char *filename)
{
  return guard([&] {
    require(cfg && vec && filename, "null argument");
    validate(*cfg);
    write_vtu_host(*cfg, make_mesh(mesh_desc(*cfg)), vec, t, filename);
  });
}

int
dg_host_index_map(const dg_config *cfg, int64_t *n_cells,
                  double *cell_centres, int64_t *dof_offsets)
{
  return guard([&] {
    require(cfg && n_cells, "null argument");
    validate(*cfg);
    const StructuredMesh m  = make_mesh(mesh_desc(*cfg));
    const int            n  = cfg->degree + 1;
    const int64_t        nd = (int64_t)(cfg->dim == 3 ? n * n * n : n * n) *
                       cfg->n_components;
    *n_cells = m.n_cells;
    const int nv = 1 << cfg->dim;
    for (int64_t c = 0; c < m.n_cells; ++c)
      {
        if (cell_centres)
          for (int d = 0; d < cfg->dim; ++d)
            {
              double s = 0;
              for (int v = 0; v < nv; ++v)
                s += m.cell_vertices[(c * nv + v) * cfg->dim + d];
              cell_centres[c * cfg->dim + d] = s / nv;
            }
        if (dof_offsets)
          dof_offsets[c] = c * nd;
      }
  });
}

} // extern "C"
