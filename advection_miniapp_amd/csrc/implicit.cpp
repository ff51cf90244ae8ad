// Implicit operators of the C ABI: the single-stage operator M/dt + L and
// its inverse mass, the Radau IIA stage-coupled operator, the cellwise FDM
// preconditioner, restarted (F)GMRES on device vectors, and the DIRK / IRK
// time steps that drive them.
#include <cmath>
#include <cstring>
#include <functional>
#include <map>

#include "aux.hpp"
#include "eigen.hpp"
#include "internal.hpp"
#include "registry.hpp"

namespace dgops
{
  std::map<uint32_t, FdmLaunch> &
  fdm_registry()
  {
    static std::map<uint32_t, FdmLaunch> m;
    return m;
  }

  namespace
  {
    thread_local std::string msg;

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

    Operator *
    get(dg_operator *h)
    {
      require(h != nullptr, "null operator handle");
      return reinterpret_cast<Operator *>(h);
    }

    // Radau IIA matrices as tabulated in advection_solver_irk_real.cc:
    // 344-407 (the 4-stage table is accurate to about 7 digits there; the
    // same values are used so that results match the reference)
    std::vector<double>
    radau_table(int s)
    {
      switch (s)
        {
          case 1:
            return {1.};
          case 2:
            return {5. / 12., -1. / 12., 3. / 4., 1. / 4.};
          case 3:
            return {0.19681547722366044, -0.06553542585019838,
                    0.02377097434822015, 0.3944243147390873,
                    0.29207341166522843, -0.04154875212599792,
                    0.37640306270046725, 0.5124858261884216,
                    1. / 9.};
          case 4:
            return {0.11299954065227377,  -0.04030922142462728,
                    0.025802391441157962, -0.009904667574594527,
                    0.2343839541547278,   0.20689257722452742,
                    -0.047857142857142855, 0.016047419401474628,
                    0.21668177697189483,  0.4061232017705644,
                    0.18903654485049834,  -0.02418210484824596,
                    0.22046204620462045,  0.3881932021466905,
                    0.3288439955106622,   1. / 16.};
          case 5:
            return {0.072998864317903,  -0.026735331107945, 0.018676929763984,
                    -0.012879106093306, 0.005042839233882,  0.153775231479182,
                    0.146214867847493,  -0.036444568905128, 0.021233063119304,
                    -0.007935579902729, 0.140063045684809,  0.298967129491283,
                    0.167585070135249,  -0.033969101686618, 0.010944288744193,
                    0.144894308109531,  0.276500068760161,  0.325797922910423,
                    0.128756753254908,  -0.015708917378805, 0.143713560791217,
                    0.281356015149465,  0.311826522975736,  0.223103901083572,
                    1. / 25.};
          default:
            throw Error(DG_ERR_UNSUPPORTED, "Radau IIA with 1..5 stages");
        }
    }

    // the quadrature switches (use_gl_quad, use_gl_quad_mass) are built for
    // the explicit operators; the implicit operators, their FDM and the IRK
    // path integrate with QGauss(k+1)
    void
    require_gauss_quadrature(const Operator &op)
    {
      require(!op.cfg.use_gl_quad && !op.cfg.use_gl_quad_mass,
              "implicit operators: built for QGauss(k+1) (use_gl_quad and "
              "use_gl_quad_mass apply to the explicit operator)",
              DG_ERR_UNSUPPORTED);
    }

    void
    setup_irk(Operator &op)
    {
      require_gauss_quadrature(op);
      require(op.cfg.operator_kind == DG_OP_ADVECTION,
              "Radau IIA operators: built for advection (the IRK solvers of "
              "advection_solver_irk_real.cc solve advection only)",
              DG_ERR_UNSUPPORTED);
      const int s = op.cfg.irk_stages;
      require(s >= 1 && s <= 4, "IRK needs 1..4 stages", DG_ERR_UNSUPPORTED);
      if (op.irk.s == s)
        return;
      IrkData &I = op.irk;
      I.s        = s;
      I.A        = radau_table(s);
      I.b.assign(s, 0.);
      I.c.assign(s, 0.);
      for (int i = 0; i < s; ++i)
        {
          I.b[i] = I.A[(s - 1) * s + i];
          for (int j = 0; j < s; ++j)
            I.c[i] += I.A[i * s + j];
        }
      I.invA           = invert_dense(I.A, s);
      const auto eig   = real_pair_eigen(I.invA, s);
      I.T              = eig.V;
      I.invT           = eig.Vinv;
      I.lam_re         = eig.lam_re;
      I.lam_im         = eig.lam_im;
      I.pair           = eig.kind;
    }

    template <typename T>
    void
    fill_spatial(FdmDev<T> &F, const RealPairEig &E,
                 const std::vector<double> &VinvMinv, int n)
    {
      for (int i = 0; i < n; ++i)
        {
          for (int j = 0; j < n; ++j)
            {
              F.VinvMinv[i][j] = (T)VinvMinv[i * n + j];
              F.V[i][j]        = (T)E.V[i * n + j];
            }
          F.lre[i] = (T)E.lam_re[i];
          F.lim[i] = (T)E.lam_im[i];
        }
      int nb = 0;
      for (int i = 0; i < n;)
        {
          F.kind[i]          = E.kind[i];
          F.block_col[nb]    = i;
          F.block_size[nb]   = E.kind[i] == 1 ? 2 : 1;
          if (E.kind[i] == 1)
            F.kind[i + 1] = 2;
          i += F.block_size[nb];
          ++nb;
        }
      F.n_blocks = nb;
    }

    template <typename T>
    void
    fill_stages(FdmDev<T> &F, const IrkData *I)
    {
      if (!I)
        {
          F.Tinv[0][0] = F.Tm[0][0] = 1;
          F.mre[0]                  = 1;
          F.mim[0]                  = 0;
          F.n_stage_blocks          = 1;
          F.stage_block_col[0]      = 0;
          F.stage_block_size[0]     = 1;
          return;
        }
      const int s = I->s;
      for (int i = 0; i < s; ++i)
        {
          for (int j = 0; j < s; ++j)
            {
              F.Tinv[i][j] = (T)I->invT[i * s + j];
              F.Tm[i][j]   = (T)I->T[i * s + j];
            }
          F.mre[i] = (T)I->lam_re[i];
          F.mim[i] = (T)I->lam_im[i];
        }
      int nb = 0;
      for (int i = 0; i < s;)
        {
          F.stage_block_col[nb]  = i;
          F.stage_block_size[nb] = I->pair[i] == 1 ? 2 : 1;
          i += F.stage_block_size[nb];
          ++nb;
        }
      F.n_stage_blocks = nb;
    }

    // the step-3 task records of one cell, sorted by their number of paired
    // factors (see FdmDev::tasks)
    template <typename T>
    void
    fill_tasks(FdmDev<T> &F, int dim)
    {
      std::vector<FdmTask<T>> by_nb[5];
      const int               nbd  = F.n_blocks;
      const int               nspb = dim == 3 ? nbd * nbd * nbd : nbd * nbd;
      for (int sb = 0; sb < F.n_stage_blocks; ++sb)
        for (int rb = 0; rb < nspb; ++rb)
          {
            const int  bd[3] = {rb % nbd, (rb / nbd) % nbd,
                                dim == 3 ? rb / (nbd * nbd) : 0};
            FdmTask<T> tk    = {};
            const int  js    = F.stage_block_col[sb];
            const bool sp    = F.stage_block_size[sb] == 2;
            require(js < 4 && F.n_blocks <= max_n, "FDM task encoding",
                    DG_ERR_UNSUPPORTED);
            tk.code  = js | (sp ? 4 : 0);
            tk.re[0] = F.mre[js];
            tk.im[0] = sp ? F.mim[js] : (T)0;
            int nb   = sp ? 1 : 0;
            for (int d = 0; d < dim; ++d)
              {
                const int  c  = F.block_col[bd[d]];
                const bool pr = F.block_size[bd[d]] == 2;
                require(c < 8, "FDM task encoding", DG_ERR_UNSUPPORTED);
                tk.code |= (c | (pr ? 8 : 0)) << (3 + 4 * d);
                tk.re[1 + d] = F.lre[c];
                tk.im[1 + d] = pr ? F.lim[c] : (T)0;
                nb += pr ? 1 : 0;
              }
            by_nb[nb].push_back(tk);
          }
      int k = 0;
      for (int nb = 0; nb <= 4; ++nb)
        {
          F.task_start[nb] = k;
          for (const auto &tk : by_nb[nb])
            {
              require(k < max_fdm_tasks, "FDM task table overflow",
                      DG_ERR_UNSUPPORTED);
              F.tasks[k++] = tk;
            }
        }
      F.task_start[5] = k;
    }

    // 1D upwind-DG matrices of the + velocity sign and their real-pair
    // eigendecomposition (advection_solver_irk_real.cc:1490-1541); the
    // - sign is its mirror image (see fdm_kernel.hpp), cell data
    void
    setup_fdm(Operator &op)
    {
      require_gauss_quadrature(op);
      if (op.d_fdm_d.ptr)
        return;
      const int      n     = op.n;
      const Basis1D &B     = op.basis;
      const double   alpha = op.cfg.flux_alpha;
      // tables [basis][stages]: basis 0 = Gauss-Lobatto nodal (the dof
      // basis), 1 = Lagrange at the Gauss points (the collocation basis of
      // the batched cell solver, advection_solver_irk_real.cc:1492-1494);
      // stages 0 = single stage, 1 = the IRK stage transform
      FdmDev<double> hd[4] = {};
      FdmDev<float>  hf[4] = {};
      if (op.cfg.irk_stages >= 1 && op.cfg.irk_stages <= 4)
        setup_irk(op);
      for (int colloc = 0; colloc < 2; ++colloc)
        {
          // K(i,j) = -int phi_i' phi_j + upwind end terms, M(i,j) = int
          // phi_i phi_j, both with QGauss(n) (:1497-1522)
          std::vector<double> K(n * n, 0.), M(n * n, 0.);
          if (!colloc)
            {
              for (int q = 0; q < n; ++q)
                for (int i = 0; i < n; ++i)
                  for (int j = 0; j < n; ++j)
                    {
                      K[i * n + j] -= B.G[q][i] * B.S[q][j] * B.gauss_wts[q];
                      M[i * n + j] += B.S[q][i] * B.S[q][j] * B.gauss_wts[q];
                    }
              // nodal basis at 0 and 1: phi_i(0) = delta_i0, phi_i(1) =
              // delta_i,n-1
              K[0] += -(0.5 - alpha * 0.5);
              K[(n - 1) * n + (n - 1)] += (0.5 + alpha * 0.5);
            }
          else
            {
              // phi_i(x_q) = delta_iq, phi_i'(x_q) = D[q][i]; at the ends
              // the Gauss-Lagrange values l0, l1
              for (int i = 0; i < n; ++i)
                {
                  M[i * n + i] = B.gauss_wts[i];
                  for (int j = 0; j < n; ++j)
                    K[i * n + j] = -B.D[j][i] * B.gauss_wts[j] -
                                   B.l0[i] * B.l0[j] * (0.5 - alpha * 0.5) +
                                   B.l1[i] * B.l1[j] * (0.5 + alpha * 0.5);
                }
            }
          const auto Minv = invert_dense(M, n);
          std::vector<double> D(n * n, 0.);
          for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
              for (int k = 0; k < n; ++k)
                D[i * n + j] += Minv[i * n + k] * K[k * n + j];
          const RealPairEig   E = real_pair_eigen(D, n);
          std::vector<double> VinvMinv(n * n, 0.);
          for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
              for (int k = 0; k < n; ++k)
                VinvMinv[i * n + j] += E.Vinv[i * n + k] * Minv[k * n + j];
          FdmDev<double> fd1 = {};
          FdmDev<float>  ff1 = {};
          fill_spatial(fd1, E, VinvMinv, n);
          fill_spatial(ff1, E, VinvMinv, n);
          fill_stages(fd1, nullptr);
          fill_stages(ff1, nullptr);
          FdmDev<double> fdS = fd1;
          FdmDev<float>  ffS = ff1;
          if (op.cfg.irk_stages >= 1 && op.cfg.irk_stages <= 4)
            {
              fill_stages(fdS, &op.irk);
              fill_stages(ffS, &op.irk);
            }
          hd[2 * colloc]     = fd1;
          hd[2 * colloc + 1] = fdS;
          hf[2 * colloc]     = ff1;
          hf[2 * colloc + 1] = ffS;
        }
      for (int i = 0; i < 4; ++i)
        {
          fill_tasks(hd[i], op.cfg.dim);
          fill_tasks(hf[i], op.cfg.dim);
        }
      op.d_fdm_d.alloc(sizeof(hd));
      op.d_fdm_f.alloc(sizeof(hf));
      hip_check(hipMemcpy(op.d_fdm_d.ptr, hd, sizeof(hd), hipMemcpyHostToDevice),
                "upload fdm");
      hip_check(hipMemcpy(op.d_fdm_f.ptr, hf, sizeof(hf), hipMemcpyHostToDevice),
                "upload fdm");
      op.d_cell_velocity.alloc(sizeof(double) * op.mesh.n_cells * op.cfg.dim);
      op.d_cell_invdet.alloc(sizeof(double) * op.mesh.n_cells);
      aux_fdm_cell_data(op, op.d_cell_velocity.as<double>(),
                        op.d_cell_invdet.as<double>());
    }

    // prec: 0 fp64, 1 fp32 vectors, 2 fp32 arithmetic on fp64 vectors
    void
    launch_fdm_variant(Operator &op, int ns, int prec, double vel_scale,
                       double inv_dt, const void *src, void *dst,
                       bool colloc = false)
    {
      setup_fdm(op);
      FdmParams P     = {};
      P.n_cells       = op.mesh.n_cells;
      P.cell_velocity = op.d_cell_velocity.as<double>();
      P.cell_invdet   = op.d_cell_invdet.as<double>();
      P.vel_scale     = vel_scale;
      P.inv_dt        = inv_dt;
      P.block         = op.n_dofs;
      auto it = fdm_registry().find(fdm_key(op.cfg.dim, op.n, ns, prec));
      require(it != fdm_registry().end(),
              "FDM variant not built: dim " + std::to_string(op.cfg.dim) +
                " degree " + std::to_string(op.n - 1) + " stages " +
                std::to_string(ns),
              DG_ERR_UNSUPPORTED);
      const char *tab = prec == 0 ?
                          (const char *)op.d_fdm_d.ptr :
                          (const char *)op.d_fdm_f.ptr;
      const size_t sz = prec == 0 ? sizeof(FdmDev<double>) :
                                    sizeof(FdmDev<float>);
      const void *tables = tab + (2 * (colloc ? 1 : 0) + (ns == 1 ? 0 : 1)) * sz;
      hip_check(it->second(P, tables, src, dst, op.stream), "fdm launch");
    }

    // implicit single-stage operator: dst = (M/dt + L(t+dt)) src
    void
    vmult_single(Operator &op, bool fp32, double t, double dt, const void *src,
                 void *dst, double scale = 1., bool with_mass = true)
    {
      require(op.cfg.operator_kind != DG_OP_BURGERS,
              "implicit operator: advection (+ diffusion) only",
              DG_ERR_STATE);
      require_gauss_quadrature(op);
      OpParams P = op.base_params(t + dt, 1);
      P.sign     = 1.;
      P.has_mass = (with_mass && dt != 0.) ? 1 : 0;
      P.mass[0][0] = dt != 0. ? 1. / dt : 0.;
      // SIPG on perturbed cells: the advection variant, then the
      // general-geometry diffusion terms added (FP64)
      const bool general = op.sipg_general();
      if (fp32)
        {
          require(!general,
                  "SIPG on perturbed cells is built in FP64 (the FP32 "
                  "variants need axis-aligned cells)",
                  DG_ERR_UNSUPPORTED);
          OpBuffers<float> io = {};
          io.src   = (const float *)src;
          io.dst   = (float *)dst;
          io.scale = (float)scale;
          launch_operator(op, P, 1, true, true, fin_integrate, epi_store, &io);
        }
      else
        {
          OpBuffers<double> io = {};
          io.src   = (const double *)src;
          io.dst   = (double *)dst;
          io.scale = scale;
          launch_operator(op, P, 1, false, true, fin_integrate, epi_store,
                          &io, general);
          if (general)
            aux_sipg_general(op, P, src, dst, scale);
        }
    }

    // stage-coupled IRK operator (A^-1 (x) M)/dt + diag_c L(t + c_c dt)
    void
    vmult_irk(Operator &op, bool fp32, double t, double dt, const void *src,
              void *dst)
    {
      setup_irk(op);
      const int s = op.irk.s;
      OpParams  P = op.base_params(t, s);
      P.sign      = 1.;
      P.has_mass  = 1;
      for (int c = 0; c < s; ++c)
        {
          P.vel_factor[c] = velocity_time_factor(op.cfg, t + op.irk.c[c] * dt);
          for (int b = 0; b < s; ++b)
            P.mass[c][b] = op.irk.invA[c * s + b] / dt;
        }
      if (s == 1)
        {
          // one stage: L at t + dt with mass 1/dt
          vmult_single(op, fp32, t, dt, src, dst);
          return;
        }
      if (fp32)
        {
          OpBuffers<float> io = {};
          io.src         = (const float *)src;
          io.dst         = (float *)dst;
          io.comp_stride = op.n_dofs;
          io.scale       = 1.f;
          launch_operator(op, P, s, true, true, fin_integrate, epi_store, &io);
        }
      else
        {
          OpBuffers<double> io = {};
          io.src         = (const double *)src;
          io.dst         = (double *)dst;
          io.comp_stride = op.n_dofs;
          io.scale       = 1.;
          launch_operator(op, P, s, false, true, fin_integrate, epi_store,
                          &io);
        }
    }

    // the cell-local stage operator in the Gauss-point collocation basis
    // (CellwiseOperator::vmult, advection_solver_irk_real.cc:545-711): the
    // stage-coupled operator of vmult_irk with the neighbour traces zero,
    // values at the Gauss points in and out
    void
    cell_operator(Operator &op, bool fp32, double t, double dt,
                  const void *src, void *dst)
    {
      setup_irk(op);
      const int s = op.irk.s;
      OpParams  P = op.base_params(t, s);
      P.sign      = 1.;
      P.has_mass  = 1;
      for (int c = 0; c < s; ++c)
        {
          P.vel_factor[c] = velocity_time_factor(op.cfg, t + op.irk.c[c] * dt);
          for (int b = 0; b < s; ++b)
            P.mass[c][b] = op.irk.invA[c * s + b] / dt;
        }
      if (fp32)
        {
          OpBuffers<float> io = {};
          io.src         = (const float *)src;
          io.dst         = (float *)dst;
          io.comp_stride = op.n_dofs;
          io.scale       = 1.f;
          launch_operator(op, P, s, true, true, fin_local, epi_store, &io);
        }
      else
        {
          OpBuffers<double> io = {};
          io.src         = (const double *)src;
          io.dst         = (double *)dst;
          io.comp_stride = op.n_dofs;
          io.scale       = 1.;
          launch_operator(op, P, s, false, true, fin_local, epi_store, &io);
        }
    }

    // The batched cell solver of the block-Jacobi preconditioner
    // (do_batched_solver with one iteration, advection_solver_irk_real.cc:
    // 2079-2126): the stage residual goes to the collocation basis, one
    // right-preconditioned GMRES step per cell with the collocation FDM
    // (x = alpha P^-1 r, alpha minimizing the cell residual), back to the
    // nodal basis.
    void
    cell_gmres_precondition(Operator &op, bool fp32, double t, double dt,
                            double vs, const void *src, void *dst)
    {
      setup_irk(op);
      const int     s   = op.irk.s;
      const int     dim = op.cfg.dim, n = op.n;
      const int64_t N   = op.n_dofs * s;
      const size_t  es  = fp32 ? 4 : 8;
      void         *rq  = op.solver_vector(60, es * N);
      void         *y   = op.solver_vector(61, es * N);
      void         *z   = op.solver_vector(62, es * N);
      const auto    tc  = colloc_registry().find(colloc_key(dim, n, fp32));
      const auto    cg =
        cell_gmres_registry().find(cell_gmres_key(dim, n, s, fp32));
      require(tc != colloc_registry().end() &&
                cg != cell_gmres_registry().end(),
              "batched cell solver not built for this dim / degree / "
              "number of stages",
              DG_ERR_UNSUPPORTED);
      hip_check(tc->second(op.mesh.n_cells, op.n_dofs, s,
                           op.basis_ptr(fp32), src, rq, op.stream),
                "collocation transform");
      launch_fdm_variant(op, s, fp32 ? 1 : 0, vs, 1. / dt, rq, y, true);
      cell_operator(op, fp32, t, dt, y, z);
      hip_check(cg->second(op.mesh.n_cells, op.n_dofs, op.basis_ptr(fp32), rq,
                           y, z, dst, op.stream),
                "cell solver update");
    }

    double
    avg_stage_factor(const Operator &op, double t, double dt)
    {
      require(op.irk.s > 0, "IRK stage table not set up", DG_ERR_STATE);
      double f = 0;
      for (int c = 0; c < op.irk.s; ++c)
        f += velocity_time_factor(op.cfg, t + op.irk.c[c] * dt);
      return f / op.irk.s;
    }

    // ---- restarted (F)GMRES on device vectors --------------------------------

    struct LinOp
    {
      std::function<void(const void *src, void *dst)> A;
      std::function<void(const void *src, void *dst)> P;
      int     prec; // vector precision
      int64_t n;    // vector length
    };

    struct KrylovResult
    {
      int    iterations;
      double initial, final;
      bool   converged;
    };

    // Arnoldi with classical Gram-Schmidt: one pass computes V^T w and
    // |w|^2. When more than half of |w|^2 cancels, a second sweep runs
    // (Kahan-Parlett "twice is enough"; one CGS sweep alone loses
    // orthogonality like eps |r0|^2 / |r_k|^2 and stalls GMRES near a
    // 1e-8 reduction); its projections come out of the subtraction pass.
    KrylovResult
    gmres(Operator &op, const LinOp &L, void *x, const void *b, int restart,
          int max_it, double abs_tol, bool flexible, size_t slot0)
    {
      const int     m   = std::max(1, restart);
      const int64_t n   = L.n;
      const size_t  es  = L.prec == DG_FP32 ? 4 : 8;
      const size_t  vb  = es * n;
      // V: m+1 vectors contiguous, Z: m vectors contiguous, tmp
      char *V   = (char *)op.solver_vector(slot0, vb * (m + 1));
      char *Z   = flexible ? (char *)op.solver_vector(slot0 + 1, vb * m) :
                             nullptr;
      char *tmp = (char *)op.solver_vector(slot0 + 2, vb);
      auto  vec = [&](char *base, int j) { return (void *)(base + vb * j); };

      KrylovResult res = {0, 0, 0, false};
      // r0 = b - A x
      L.A(x, tmp);
      vec_copy_convert(op, n, L.prec, b, L.prec, vec(V, 0));
      vec_axpby(op, L.prec, n, -1., tmp, 1., vec(V, 0));
      double beta2;
      vec_multi_dot(op, L.prec, n, 1, vec(V, 0), n, vec(V, 0), &beta2);
      double beta = std::sqrt(std::max(beta2, 0.));
      res.initial = res.final = beta;
      if (beta <= abs_tol)
        {
          res.converged = true;
          return res;
        }
      std::vector<double> H((m + 1) * m, 0.), cs(m), sn(m), g(m + 1);
      std::vector<double> h(m + 2), h2(m + 2);
      int total = 0;
      while (true)
        {
          if (beta == 0.)
            {
              res.converged = true;
              break;
            }
          vec_axpby(op, L.prec, n, 1. / beta, vec(V, 0), 0., vec(V, 0));
          std::fill(g.begin(), g.end(), 0.);
          g[0]  = beta;
          int j = 0;
          bool done = false;
          for (; j < m; ++j)
            {
              void *zj = flexible ? vec(Z, j) : tmp;
              L.P(vec(V, j), zj);
              void *w = vec(V, j + 1);
              L.A(zj, w);
              // h = V_{0..j}^T w and |w|^2 in one pass (x_k = V_k, k <= j+1)
              vec_multi_dot(op, L.prec, n, j + 2, V, n, w, h.data());
              const double ww = h[j + 1];
              double       hh = 0;
              for (int k = 0; k <= j; ++k)
                hh += h[k] * h[k];
              std::vector<double> neg(j + 1);
              for (int k = 0; k <= j; ++k)
                neg[k] = -h[k];
              double hn2 = ww - hh;
              if (hn2 < 0.5 * ww)
                {
                  // w -= V h with h2 = V^T w, |w|^2 of the result
                  vec_multi_axpy_dot(op, L.prec, n, j + 1, neg.data(), V, n,
                                     w, h2.data());
                  for (int k = 0; k <= j; ++k)
                    {
                      neg[k] = -h2[k];
                      h[k] += h2[k];
                    }
                  double s2 = 0;
                  for (int k = 0; k <= j; ++k)
                    s2 += h2[k] * h2[k];
                  hn2 = h2[j + 1] - s2;
                }
              // the last subtraction and the normalization in one pass:
              // |w - V c|^2 = |w|^2 - |c|^2 is known before it
              const double hn = std::sqrt(std::max(hn2, 0.));
              vec_multi_axpy_scaled(op, L.prec, n, j + 1, neg.data(), V, n, w,
                                    hn > 0 ? 1. / hn : 1.);
              for (int k = 0; k <= j; ++k)
                H[k * m + j] = h[k];
              H[(j + 1) * m + j] = hn;
              // Givens rotations
              for (int k = 0; k < j; ++k)
                {
                  const double a = H[k * m + j], bb = H[(k + 1) * m + j];
                  H[k * m + j]       = cs[k] * a + sn[k] * bb;
                  H[(k + 1) * m + j] = -sn[k] * a + cs[k] * bb;
                }
              const double a = H[j * m + j], bb = H[(j + 1) * m + j];
              const double r = std::sqrt(a * a + bb * bb);
              cs[j] = r == 0 ? 1. : a / r;
              sn[j] = r == 0 ? 0. : bb / r;
              H[j * m + j]       = r;
              H[(j + 1) * m + j] = 0.;
              g[j + 1] = -sn[j] * g[j];
              g[j]     = cs[j] * g[j];
              ++total;
              res.final = std::fabs(g[j + 1]);
              if (res.final <= abs_tol || total >= max_it || hn == 0.)
                {
                  ++j;
                  done = true;
                  break;
                }
            }
          // y = R^-1 g
          std::vector<double> y(j, 0.);
          for (int k = j - 1; k >= 0; --k)
            {
              double s = g[k];
              for (int l = k + 1; l < j; ++l)
                s -= H[k * m + l] * y[l];
              y[k] = H[k * m + k] != 0 ? s / H[k * m + k] : 0.;
            }
          if (flexible)
            vec_multi_axpy(op, L.prec, n, j, y.data(), Z, n, x);
          else
            {
              // x += P (V y)
              vec_axpby(op, L.prec, n, 0., vec(V, 0), 0., tmp);
              vec_multi_axpy(op, L.prec, n, j, y.data(), V, n, tmp);
              void *pz = vec(V, m); // free slot after the cycle
              L.P(tmp, pz);
              vec_axpby(op, L.prec, n, 1., pz, 1., x);
            }
          if (done || total >= max_it)
            {
              res.converged = res.final <= abs_tol;
              break;
            }
          // restart: r = b - A x
          L.A(x, tmp);
          vec_copy_convert(op, n, L.prec, b, L.prec, vec(V, 0));
          vec_axpby(op, L.prec, n, -1., tmp, 1., vec(V, 0));
          vec_multi_dot(op, L.prec, n, 1, vec(V, 0), n, vec(V, 0), &beta2);
          beta      = std::sqrt(std::max(beta2, 0.));
          res.final = beta;
          if (beta <= abs_tol)
            {
              res.converged = true;
              break;
            }
        }
      res.iterations = total;
      return res;
    }

    // The inner solver of the mixed-precision method: m steps of
    // right-preconditioned GMRES from x0 = 0 in the vector precision of L,
    // in flexible form (z_j = P v_j kept, x = Z y, so no preconditioner
    // application after the last step), one classical Gram-Schmidt sweep
    // per step. Every scalar stays in device memory: the whole solve is a
    // sequence of launches without a host round trip or a convergence
    // test. b and x are FP64 vectors, converted on the way in and out.
    void
    gmres_inner_device(Operator &op, const LinOp &L, void *x, const void *b,
                       int m, size_t slot0)
    {
      require(m >= 1 && m <= 16, "inner GMRES: 1..16 iterations",
              DG_ERR_UNSUPPORTED);
      const int64_t n  = L.n;
      const size_t  vb = (L.prec == DG_FP32 ? 4 : 8) * n;
      char   *V = (char *)op.solver_vector(slot0, vb * (m + 1));
      char   *Z = (char *)op.solver_vector(slot0 + 1, vb * m);
      double *S = (double *)op.solver_vector(
        slot0 + 2, sizeof(double) * (m * (m + 2) + 1 + m));
      const int64_t hs    = m + 2; // column j: H_0j .. H_jj, <w_j, w_j>
      double       *beta2 = S + m * hs;
      double       *y     = beta2 + 1;
      auto vec = [&](char *base, int j) { return (void *)(base + vb * j); };

      vec_copy_convert(op, n, DG_FP64, b, L.prec, vec(V, 0));
      vec_multi_dot_dev(op, L.prec, n, 1, vec(V, 0), n, vec(V, 0), beta2);
      vec_scale_inv_sqrt(op, L.prec, n, beta2, vec(V, 0));
      for (int j = 0; j < m; ++j)
        {
          L.P(vec(V, j), vec(Z, j));
          L.A(vec(Z, j), vec(V, j + 1));
          vec_multi_dot_dev(op, L.prec, n, j + 2, V, n, vec(V, j + 1),
                            S + j * hs);
          // the last Krylov vector is not used again: its norm enters
          // the least-squares problem only
          if (j + 1 < m)
            vec_arnoldi_update(op, L.prec, n, j + 1, S + j * hs, V, n,
                               vec(V, j + 1));
        }
      vec_arnoldi_lsq(op, m, S, hs, beta2, y);
      vec_combine(op, L.prec, n, m, y, Z, n, (double *)x);
    }
  } // namespace
} // namespace dgops

using namespace dgops;

extern "C" {

int
dg_set_time(dg_operator *h, double t, double dt)
{
  return guard([&] {
    Operator *op  = get(h);
    op->time      = t;
    op->time_step = dt;
  });
}

int
dg_vmult(dg_operator *h, void *dst, const void *src)
{
  return guard([&] {
    Operator *op = get(h);
    require(dst && src && dst != src, "vmult: distinct non-null vectors");
    vmult_single(*op, op->fp32(), op->time, op->time_step, src, dst);
  });
}

int
dg_apply_inverse_mass_matrix(dg_operator *h, void *dst)
{
  return guard([&] {
    Operator *op = get(h);
    require(dst != nullptr, "null vector");
    require(op->cfg.n_components == 1, "scalar fields only");
    aux_inverse_mass(*op, op->cfg.precision, dst);
  });
}

int
dg_precondition_block_jacobi(dg_operator *h, void *dst, const void *src)
{
  return guard([&] {
    Operator *op = get(h);
    require(dst && src, "null vector");
    require(op->time_step != 0., "precondition: set_time with dt != 0 first",
            DG_ERR_STATE);
    launch_fdm_variant(*op, 1, op->fp32() ? 1 : 0,
                       velocity_time_factor(op->cfg, op->time + op->time_step),
                       1. / op->time_step, src, dst);
  });
}

int
dg_irk_coefficients(int s, double *A, double *b, double *c)
{
  return guard([&] {
    require(A && b && c, "null output");
    const auto T = radau_table(s);
    for (int i = 0; i < s * s; ++i)
      A[i] = T[i];
    for (int i = 0; i < s; ++i)
      {
        b[i] = T[(s - 1) * s + i];
        c[i] = 0;
        for (int j = 0; j < s; ++j)
          c[i] += T[i * s + j];
      }
  });
}

int
dg_irk_compute_rhs(dg_operator *h, void *dst, const void *src)
{
  return guard([&] {
    Operator *op = get(h);
    require(dst && src, "null vector");
    require(op->time_step != 0., "compute_rhs: set_time with dt != 0",
            DG_ERR_STATE);
    setup_irk(*op);
    const size_t es = op->scalar_size();
    for (int c = 0; c < op->irk.s; ++c)
      {
        // dst_c = -1/dt L(t + c_c dt) src; vmult_single evaluates L at
        // (t' + dt'), so pass t' = t + c dt, dt' = 0 and no mass
        void *dc = (char *)dst + es * op->n_dofs * c;
        vmult_single(*op, op->fp32(), op->time + op->irk.c[c] * op->time_step,
                     0., src, dc, -1. / op->time_step, false);
      }
  });
}

int
dg_irk_vmult(dg_operator *h, void *dst, const void *src)
{
  return guard([&] {
    Operator *op = get(h);
    require(dst && src && dst != src, "distinct non-null vectors");
    require(op->time_step != 0., "irk vmult: set_time with dt != 0",
            DG_ERR_STATE);
    vmult_irk(*op, op->fp32(), op->time, op->time_step, src, dst);
  });
}

int
dg_irk_update_solution(dg_operator *h, void *solution, const void *stage)
{
  return guard([&] {
    Operator *op = get(h);
    require(solution && stage, "null vector");
    setup_irk(*op);
    const int           s = op->irk.s;
    std::vector<double> w(s, 0.);
    for (int b = 0; b < s; ++b)
      for (int c = 0; c < s; ++c)
        w[b] += op->time_step * op->irk.b[c] * op->irk.invA[c * s + b];
    vec_multi_axpy(*op, op->cfg.precision, op->n_dofs, s, w.data(), stage,
                   op->n_dofs, solution);
  });
}

int
dg_irk_precondition_block_jacobi(dg_operator *h, void *dst, const void *src)
{
  return guard([&] {
    Operator *op = get(h);
    require(dst && src, "null vector");
    require(op->time_step != 0., "precondition: set_time with dt != 0",
            DG_ERR_STATE);
    setup_irk(*op);
    launch_fdm_variant(*op, op->irk.s, op->fp32() ? 1 : 0,
                       avg_stage_factor(*op, op->time, op->time_step),
                       1. / op->time_step, src, dst);
  });
}

int
dg_irk_precondition_cell_solver(dg_operator *h, void *dst, const void *src)
{
  return guard([&] {
    Operator *op = get(h);
    require(dst && src, "null vector");
    require(op->time_step != 0., "precondition: set_time with dt != 0",
            DG_ERR_STATE);
    // the stage count enters the averaged speed factor
    setup_irk(*op);
    cell_gmres_precondition(*op, op->fp32(), op->time, op->time_step,
                            avg_stage_factor(*op, op->time, op->time_step),
                            src, dst);
  });
}

int
dg_irk_cell_operator(dg_operator *h, void *dst, const void *src)
{
  return guard([&] {
    Operator *op = get(h);
    require(dst && src, "null vector");
    require(op->time_step != 0., "cell operator: set_time with dt != 0",
            DG_ERR_STATE);
    cell_operator(*op, op->fp32(), op->time, op->time_step, src, dst);
  });
}

} // extern "C"

namespace
{
  // The preconditioner the outer (F)GMRES of the IRK solve applies, as
  // selected by a solver control: the FDM block-Jacobi step (or the batched
  // cell solver) in the control's precision, or an inner GMRES of a fixed
  // number of steps around it (BlockJacobi::vmult,
  // advection_solver_irk_real.cc:2462-2479, with the build's precision and
  // inner-solver choices). Vectors are FP64 block vectors.
  std::function<void(const void *, void *)>
  irk_outer_preconditioner(Operator &op, const dg_solver_control &ctl_in)
  {
    require(ctl_in.inner_iterations >= 0 && ctl_in.inner_iterations <= 16,
            "inner GMRES: 0..16 iterations", DG_ERR_UNSUPPORTED);
    setup_irk(op);
    const int     s     = op.irk.s;
    const int64_t N     = op.n_dofs * s;
    const double  t     = op.time, dt = op.time_step;
    const double  vs    = avg_stage_factor(op, t, dt);
    const int     pp    = ctl_in.prec_precision;
    const int     inner = ctl_in.inner_iterations;
    const bool    cs    = ctl_in.cell_solver != 0;
    Operator     *o     = &op;
    if (inner > 0)
      {
        // inner right-preconditioned GMRES in `pp` with the FDM (or the
        // batched cell solver), a fixed number of steps on the device
        require(ctl_in.flexible, "an inner solver needs the flexible method");
        const bool f32 = pp == DG_FP32;
        return [=](const void *src, void *dst) {
          LinOp in;
          in.prec = f32 ? DG_FP32 : DG_FP64;
          in.n    = N;
          in.A    = [=](const void *x, void *y) {
            vmult_irk(*o, f32, t, dt, x, y);
          };
          in.P = [=](const void *x, void *y) {
            if (cs)
              cell_gmres_precondition(*o, f32, t, dt, vs, x, y);
            else
              launch_fdm_variant(*o, s, f32 ? 1 : 0, vs, 1. / dt, x, y);
          };
          gmres_inner_device(*o, in, dst, src, inner, 20);
        };
      }
    const int prec = pp == DG_FP32 ? 2 : 0;
    require(!cs || prec == 0,
            "the batched cell solver runs in the precision of its "
            "vectors: FP32 through an inner solver (inner_iterations)",
            DG_ERR_UNSUPPORTED);
    return [=](const void *x, void *y) {
      if (cs)
        cell_gmres_precondition(*o, false, t, dt, vs, x, y);
      else
        launch_fdm_variant(*o, s, prec, vs, 1. / dt, x, y);
    };
  }
} // namespace

extern "C" {

int
dg_irk_block_jacobi_vmult(dg_operator *h, const dg_solver_control *ctl,
                          void *dst, const void *src)
{
  return guard([&] {
    Operator *op = get(h);
    require(dst && src && ctl && dst != src,
            "distinct non-null vectors and a control");
    require(!op->fp32(), "the preconditioner of the FP64 outer solver "
                         "takes FP64 vectors: create the operator with "
                         "precision DG_FP64",
            DG_ERR_STATE);
    require(op->time_step != 0., "preconditioner: set_time with dt != 0",
            DG_ERR_STATE);
    irk_outer_preconditioner(*op, *ctl)(src, dst);
  });
}

int
dg_irk_solve(dg_operator *h, void *stage, const void *rhs,
             dg_solver_control *ctl)
{
  return guard([&] {
    Operator *op = get(h);
    require(stage && rhs && ctl, "null argument");
    require(!op->fp32(), "the outer solver runs in FP64: create the "
                         "operator with precision DG_FP64",
            DG_ERR_STATE);
    require(op->time_step != 0., "irk solve: set_time with dt != 0",
            DG_ERR_STATE);
    setup_irk(*op);
    const int     s  = op->irk.s;
    const int64_t N  = op->n_dofs * s;
    const double  t  = op->time, dt = op->time_step;
    LinOp outer;
    outer.prec = DG_FP64;
    outer.n    = N;
    outer.A    = [&](const void *x, void *y) {
      vmult_irk(*op, false, t, dt, x, y);
    };
    outer.P = irk_outer_preconditioner(*op, *ctl);
    double bn2;
    vec_multi_dot(*op, DG_FP64, N, 1, rhs, N, rhs, &bn2);
    const double tol = std::max(ctl->abs_tolerance,
                                ctl->rel_tolerance * std::sqrt(bn2));
    const auto r = gmres(*op, outer, stage, rhs, ctl->restart,
                         ctl->max_iterations, tol, ctl->flexible != 0, 0);
    ctl->iterations       = r.iterations;
    ctl->initial_residual = r.initial;
    ctl->final_residual   = r.final;
    if (!r.converged)
      throw Error(DG_ERR_NO_CONVERGENCE,
                  "IRK solve did not converge in " +
                    std::to_string(r.iterations) + " iterations, residual " +
                    std::to_string(r.final));
  });
}

int
dg_irk_time_step(dg_operator *h, double t, double dt, void *solution,
                 dg_solver_control *ctl)
{
  return guard([&] {
    Operator *op = get(h);
    require(solution && ctl, "null argument");
    op->time      = t;
    op->time_step = dt;
    setup_irk(*op);
    const int     s   = op->irk.s;
    const int64_t N   = op->n_dofs * s;
    void         *rhs = op->solver_vector(30, 8 * N);
    void         *stg = op->solver_vector(31, 8 * N);
    int rc = dg_irk_compute_rhs(h, rhs, solution);
    if (rc != DG_OK)
      throw Error(rc, dg_last_error());
    vec_axpby(*op, DG_FP64, N, 0., rhs, 0., stg);
    dg_solver_control c = *ctl;
    rc = dg_irk_solve(h, stg, rhs, &c);
    ctl->iterations       = c.iterations;
    ctl->initial_residual = c.initial_residual;
    ctl->final_residual   = c.final_residual;
    if (rc != DG_OK)
      throw Error(rc, dg_last_error());
    rc = dg_irk_update_solution(h, solution, stg);
    if (rc != DG_OK)
      throw Error(rc, dg_last_error());
  });
}

int
dg_dirk_time_step(dg_operator *h, int n_stages, double time, double dt,
                  void *solution, dg_solver_control *ctl)
{
  // DiagonallyImplicitRungeKuttaIntegrator::perform_time_step
  // (advection_solver_implicit_real.cc:1653-1731) on device vectors
  return guard([&] {
    Operator *op = get(h);
    require(solution && ctl, "null argument");
    const int  s   = n_stages;
    std::vector<double> b(s, 0.), A(s * s, 0.), c(s, 0.);
    if (s == 1)
      b[0] = 1.;
    else if (s == 3)
      {
        const double gamma = 0.4358665215084589994160194;
        const double alpha = 1 + gamma * (-4 + 2 * gamma);
        const double beta  = -1 + gamma * (6 + gamma * (-9 + 3 * gamma));
        b                  = {(-1 + 4 * gamma) / (4 * beta),
                              -0.75 * alpha * alpha / beta, gamma};
        for (int d = 0; d < 3; ++d)
          A[d * s + d] = gamma;
        const double c2 = (2 + gamma * (-9 + 6 * gamma)) / (3 * alpha);
        A[1 * s + 0]    = c2 - gamma;
      }
    else if (s == 4)
      {
        const double gamma = 9. / 40.;
        b = {4032. / 9943., 6929. / 15485., -723. / 9272., gamma};
        for (int d = 0; d < 4; ++d)
          A[d * s + d] = gamma;
        A[1 * s + 0] = 163. / 520.;
        A[2 * s + 0] = -6481433. / 8838675.;
        A[2 * s + 1] = 87795409. / 70709400.;
      }
    else if (s == 7)
      {
        const double gamma = 1. / 8.;
        b = {-5649241495537. / 14093099002237., -5649241495537. / 14093099002237.,
             5718691255176. / 6089204655961.,   2199600963556. / 4241893152925.,
             8860614275765. / 11425531467341.,  -3696041814078. / 6641566663007.,
             gamma};
        A[0] = 0.;
        for (int d = 1; d < s; ++d)
          A[d * s + d] = gamma;
        A[2 * s + 1] = -39188347878. / 1513744654945.;
        A[3 * s + 1] = 1748874742213. / 5168247530883.;
        A[3 * s + 2] = -1748874742213. / 5795261096931.;
        A[4 * s + 1] = -6429340993097. / 17896796106705.;
        A[4 * s + 2] = 9711656375562. / 10370074603625.;
        A[4 * s + 3] = 1137589605079. / 3216875020685.;
        A[5 * s + 1] = 405169606099. / 1734380148729.;
        A[5 * s + 2] = -264468840649. / 6105657584947.;
        A[5 * s + 3] = 118647369377. / 6233854714037.;
        A[5 * s + 4] = 683008737625. / 4934655825458.;
        for (int d = 1; d < 6; ++d)
          A[d * s + 0] = A[d * s + 1];
      }
    else
      throw Error(DG_ERR_UNSUPPORTED,
                  "A scheme with " + std::to_string(s) + " is not implemented!");
    for (int d = 0; d < s; ++d)
      A[(s - 1) * s + d] = b[d];
    for (int i = 0; i < s; ++i)
      for (int j = 0; j <= i; ++j)
        c[i] += A[i * s + j];

    const int     p  = op->cfg.precision;
    const size_t  es = op->scalar_size();
    const int64_t n  = op->n_dofs;
    std::vector<void *> ki(s);
    for (int i = 0; i < s; ++i)
      ki[i] = op->solver_vector(40 + i, es * n);
    void *tmp  = op->solver_vector(50, es * n);
    void *tmp2 = op->solver_vector(51, es * n);
    vec_copy_convert(*op, n, p, solution, p, tmp2);
    int          first = 0;
    const double gamma = A[(s - 1) * s + (s - 1)];
    LinOp        L;
    L.prec = p;
    L.n    = n;
    L.A    = [&](const void *x, void *y) {
      vmult_single(*op, op->fp32(), op->time, op->time_step, x, y);
    };
    L.P = [&](const void *x, void *y) {
      launch_fdm_variant(*op, 1, op->fp32() ? 1 : 0,
                         velocity_time_factor(op->cfg,
                                              op->time + op->time_step),
                         1. / op->time_step, x, y);
    };
    int total_it = 0;
    if (A[0] == 0.0)
      {
        first = 1;
        op->time      = time;
        op->time_step = 0.;
        vmult_single(*op, op->fp32(), time, 0., solution, ki[0]);
        aux_inverse_mass(*op, p, ki[0]);
        vec_axpby(*op, p, n, -b[0] * dt, ki[0], 1., solution);
        vec_axpby(*op, p, n, 0., ki[0], gamma * dt, ki[0]);
      }
    for (int stage = first; stage < s; ++stage)
      {
        op->time      = time + c[stage] * dt;
        op->time_step = 0.;
        for (int r = 1; r < stage; ++r)
          vec_axpby(*op, p, n,
                    (A[(stage - 1) * s + r - 1] - A[stage * s + r - 1]) / gamma,
                    ki[r - 1], 1., tmp2);
        if (stage > 0)
          vec_axpby(*op, p, n, -A[stage * s + stage - 1] / gamma,
                    ki[stage - 1], 1., tmp2);
        vmult_single(*op, op->fp32(), op->time, 0., tmp2, tmp);
        op->time_step = gamma * dt;
        double tn2;
        vec_multi_dot(*op, p, n, 1, tmp, n, tmp, &tn2);
        const double tol = std::max(ctl->abs_tolerance,
                                    std::sqrt(tn2) * (ctl->rel_tolerance > 0 ?
                                                        ctl->rel_tolerance :
                                                        1e-8));
        vec_axpby(*op, p, n, 0., tmp, 0., ki[stage]);
        const bool flexible = ctl->flexible || op->fp32();
        const auto r = gmres(*op, L, ki[stage], tmp,
                             ctl->restart > 0 ? ctl->restart : 10,
                             ctl->max_iterations > 0 ? ctl->max_iterations :
                                                       1000,
                             tol, flexible, 60);
        total_it += r.iterations;
        if (!r.converged)
          throw Error(DG_ERR_NO_CONVERGENCE, "DIRK stage solve did not converge");
        vec_axpby(*op, p, n, -b[stage] / gamma, ki[stage], 1., solution);
      }
    ctl->iterations = total_it;
  });
}

} // extern "C"

// ---- Radau IIA integrator with extrapolated initial guesses ----------------

namespace dgops
{
  struct IrkIntegrator
  {
    Operator *op;
    int       s;
    int64_t   N; // s * n_dofs
    DevBuf    rhs;
    DevBuf    sol[4], mv[4];
    int       sol_idx[4] = {0, 1, 2, 3}; // stage_sol[k] -> buffer
    int       mv_idx[4]  = {0, 1, 2, 3};
    unsigned  timestep_number      = 1;
    double    last_refresh_quality = 1;
    unsigned  steps_since_refresh  = 0;
    bool      refresh              = true;
    double    project_sol[4]       = {0, 0, 0, 0};

    double *stage_sol(int k) { return sol[sol_idx[k]].as<double>(); }
    double *stage_mv(int k) { return mv[mv_idx[k]].as<double>(); }
  };
} // namespace dgops

extern "C" {

int
dg_irk_integrator_create(dg_operator *h, dg_irk_integrator **out)
{
  return guard([&] {
    Operator *op = get(h);
    require(out != nullptr, "null output");
    require(!op->fp32(), "IRK integrator: FP64 operator", DG_ERR_STATE);
    setup_irk(*op);
    auto it = std::make_unique<IrkIntegrator>();
    it->op  = op;
    it->s   = op->irk.s;
    it->N   = op->n_dofs * it->s;
    it->rhs.alloc(8 * it->N);
    for (int k = 0; k < 4; ++k)
      {
        it->sol[k].alloc(8 * it->N);
        it->mv[k].alloc(8 * it->N);
        hip_check(hipMemsetAsync(it->sol[k].ptr, 0, 8 * it->N, op->stream),
                  "memset");
        hip_check(hipMemsetAsync(it->mv[k].ptr, 0, 8 * it->N, op->stream),
                  "memset");
      }
    *out = reinterpret_cast<dg_irk_integrator *>(it.release());
  });
}

int
dg_irk_integrator_destroy(dg_irk_integrator *it)
{
  return guard([&] { delete reinterpret_cast<IrkIntegrator *>(it); });
}

int
dg_irk_integrator_step(dg_irk_integrator *hi, double time, double dt,
                       void *solution, dg_solver_control *ctl,
                       double *project_out)
{
  return guard([&] {
    require(hi && solution && ctl, "null argument");
    IrkIntegrator &I  = *reinterpret_cast<IrkIntegrator *>(hi);
    Operator      &op = *I.op;
    dg_operator   *h  = reinterpret_cast<dg_operator *>(&op);
    op.time           = time;
    op.time_step      = dt;
    int rc = dg_irk_compute_rhs(h, I.rhs.ptr, solution);
    if (rc != DG_OK)
      throw Error(rc, dg_last_error());
    constexpr int nv = 4;
    if (I.refresh)
      {
        // least-squares fit of rhs by A x_{n-j} (advection_solver_irk_real.cc
        // :2544-2695)
        const unsigned n_max =
          I.timestep_number > nv ? nv : I.timestep_number - 1;
        for (unsigned st = 0; st < n_max; ++st)
          vmult_irk(op, false, time, dt, I.stage_sol(st), I.stage_mv(st));
        const double *v[4] = {I.stage_mv(0), I.stage_mv(1), I.stage_mv(2),
                              I.stage_mv(3)};
        double sums[14];
        vec_gram4(op, I.N, v, I.rhs.as<double>(), sums);
        double P[nv][nv] = {};
        for (int i = 0, c = 0; i < nv; ++i)
          for (int j = 0; j <= i; ++j, ++c)
            P[i][j] = sums[c];
        int i = 0;
        for (; i < nv; ++i)
          {
            for (int j = 0; j < i; ++j)
              {
                const double inv_entry = P[i][j] / P[j][j];
                for (int k = j + 1; k <= i; ++k)
                  P[i][k] -= P[k][j] * inv_entry;
              }
            if (P[i][i] < 1e-12 * P[0][0] || P[0][0] < 1e-30)
              break;
          }
        for (int st = 0; st < i; ++st)
          {
            I.project_sol[st] = sums[nv * (nv + 1) / 2 + st];
            for (int j = 0; j < st; ++j)
              I.project_sol[st] -= P[st][j] / P[j][j] * I.project_sol[j];
          }
        for (int st = i; st < nv; ++st)
          I.project_sol[st] = 0.;
        for (int st = i - 1; st >= 0; --st)
          {
            double sum = I.project_sol[st];
            for (int j = st + 1; j < i; ++j)
              sum -= I.project_sol[j] * P[j][st];
            I.project_sol[st] = sum / P[st][st];
          }
      }
    // extrapolate into the oldest slot, then rotate it to the front
    {
      double *dst3 = I.stage_sol(3);
      // dst3 = p3 x3 + p0 x0 + p1 x1 + p2 x2 (x3 is read before written)
      vec_axpby(op, DG_FP64, I.N, 0., dst3, I.project_sol[3], dst3);
      const double w[3] = {I.project_sol[0], I.project_sol[1],
                           I.project_sol[2]};
      for (int k = 0; k < 3; ++k)
        vec_axpby(op, DG_FP64, I.N, w[k], I.stage_sol(k), 1., dst3);
      std::swap(I.sol_idx[2], I.sol_idx[3]);
      std::swap(I.sol_idx[1], I.sol_idx[2]);
      std::swap(I.sol_idx[0], I.sol_idx[1]);
    }
    // solve from the extrapolated guess, SolverControl(max, tol * |rhs|)
    dg_solver_control c = *ctl;
    rc = dg_irk_solve(h, I.stage_sol(0), I.rhs.ptr, &c);
    ctl->iterations       = c.iterations;
    ctl->initial_residual = c.initial_residual;
    ctl->final_residual   = c.final_residual;
    if (rc != DG_OK)
      throw Error(rc, dg_last_error());
    // refresh heuristic (:2758-2773)
    if (I.refresh)
      {
        I.last_refresh_quality = c.initial_residual;
        if (I.timestep_number > 4)
          {
            I.refresh             = false;
            I.steps_since_refresh = 0;
          }
      }
    else
      {
        if (I.last_refresh_quality < 0.5 * c.initial_residual ||
            I.steps_since_refresh > 20)
          I.refresh = true;
        ++I.steps_since_refresh;
      }
    rc = dg_irk_update_solution(h, solution, I.stage_sol(0));
    if (rc != DG_OK)
      throw Error(rc, dg_last_error());
    ++I.timestep_number;
    if (project_out)
      for (int k = 0; k < 4; ++k)
        project_out[k] = I.project_sol[k];
  });
}

} // extern "C"
