// Host-side state of an operator handle, shared by the host translation
// units of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dgops.h"
#include "basis.hpp"
#include "dev_basis.hpp"
#include "dg_common.hpp"
#include "mesh.hpp"

namespace dgops
{
  struct Error : std::runtime_error
  {
    int code;
    Error(int c, const std::string &m)
      : std::runtime_error(m)
      , code(c)
    {}
  };

  inline void
  hip_check(hipError_t e, const char *what)
  {
    if (e != hipSuccess)
      throw Error(DG_ERR_HIP,
                  std::string(what) + ": " + hipGetErrorString(e));
  }

  void set_last_error(const std::string &msg);

  // device buffer owned by a handle
  struct DevBuf
  {
    void  *ptr   = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
      if (ptr)
        (void)hipFree(ptr);
      ptr   = nullptr;
      bytes = 0;
    }
    void alloc(size_t b)
    {
      if (b <= bytes && ptr)
        return;
      release();
      hip_check(hipMalloc(&ptr, b ? b : 8), "hipMalloc");
      bytes = b;
    }
    template <typename T>
    T *as() const
    {
      return static_cast<T *>(ptr);
    }
  };

  // FDM data of one cell-independent 1D problem (advection sign c)
  struct FdmTables1D
  {
    // real-pair eigenbasis: D = M^-1 K, D V = V Lambda (block diagonal)
    double V[max_n][max_n];     // [i][j]
    double VinvMinv[max_n][max_n];
    double lambda_re[max_n], lambda_im[max_n]; // per column of V; pairs
    int    pair[max_n];         // -1: real, else index of partner column
  };

  struct IrkData
  {
    int                 s = 0;
    std::vector<double> A, invA, b, c; // s*s, s*s, s, s
    // real-pair eigen decomposition invA = T Lambda T^-1
    std::vector<double> T, invT, lam_re, lam_im;
    std::vector<int>    pair;
  };

  // tile shape of an operator launch (autotune.cpp): shape index (cells
  // per workgroup, shape_cells in op_kernel.hpp) and the tuner's time of
  // that candidate
  struct TileChoice
  {
    int    shape = 0;
    double ms    = 0;
    bool   tuned = false;
  };

  struct Operator
  {
    dg_config      cfg;
    StructuredMesh mesh;
    int            n;      // points per direction
    int            nd;     // dofs per cell per component
    int64_t        n_dofs; // per component
    hipStream_t    stream = nullptr;
    int            device = 0;

    Basis1D basis;
    DevBuf  d_basis_d, d_basis_f, d_neighbors, d_vertices;
    // use_gl_quad_mass without use_gl_quad: the mass matrix is integrated
    // with QGaussLobatto(k+1) (its basis in d_mbasis_*), the operator with
    // QGauss(k+1). The apply then runs in two passes; stages keep their
    // right-hand side in d_gl_k (also for SIPG on perturbed cells)
    bool   gl_mass_only = false;
    DevBuf d_mbasis_d, d_mbasis_f, d_gl_k;
    // SIPG on perturbed cells: face geometry of the face above every cell
    // in every direction, built on first use (sipg.hip)
    DevBuf d_sipg_faces;
    bool   sipg_faces_ready = false;

    // implicit state
    double time      = 0;
    double time_step = 0;
    IrkData irk;
    FdmTables1D fdm[2]; // velocity sign + and -
    DevBuf  d_fdm_d, d_fdm_f, d_cell_velocity, d_cell_invdet;
    DevBuf  d_burgers; // over-integration tables of the Burgers operator

    // tabulated geometry of the operator kernels ([0] fp64, [1] fp32):
    // built on first use, time independent (the velocity's time factor
    // is applied in the kernel)
    DevBuf d_tab_cell[2], d_tab_face[2];
    bool   tables_ready[2] = {false, false};
    // boundary faces of every geometry mode: index per cell face, a . n dA
    // and x per face point (built once), Dirichlet data at those points
    // (refreshed per launch)
    DevBuf  d_tab_bnd, d_bnd_index, d_bnd_value;
    bool    boundary_ready = false;
    int64_t n_bnd          = 0; // boundary faces

    // tile shapes: forced by dg_set_tile_shape (0 = autotuned), and the
    // choice of the last operator launch with its variant key
    int        tile_force_cells = 0;
    TileChoice last_tile;
    uint32_t   last_tile_key = 0;

    // captured HIP graph of a run of LSRK steps (dg_lsrk_time_steps) and
    // what it was captured for: scheme, time step and the three registers;
    // `mode` is the launch form the first long call measured faster
    // (0 not measured, 1 stream launches, 2 graph replay) with both times,
    // `force` the caller's choice (0 = measure)
    struct StepGraph
    {
      hipGraphExec_t exec = nullptr;
      int            scheme = -1;
      double         dt     = 0;
      void          *reg[3] = {nullptr, nullptr, nullptr};
      int            force  = 0;
      int            mode   = 0;
      float          eager_ms = 0.f, graph_ms = 0.f;
      ~StepGraph()
      {
        if (exec)
          (void)hipGraphExecDestroy(exec);
      }
    } step_graph;

    // scratch for reductions and solvers
    DevBuf d_partials, d_result;
    std::vector<DevBuf *> scratch_vectors;
    std::vector<std::unique_ptr<DevBuf>> solver_vectors;

    OpParams base_params(double time, int n_comp) const;
    const void *basis_ptr(bool fp32) const
    {
      return fp32 ? d_basis_f.ptr : d_basis_d.ptr;
    }
    // basis of the mass-matrix quadrature (inverse mass)
    const void *mass_basis_ptr(bool fp32) const
    {
      if (!gl_mass_only)
        return basis_ptr(fp32);
      return fp32 ? d_mbasis_f.ptr : d_mbasis_d.ptr;
    }
    bool fp32() const { return cfg.precision == DG_FP32; }
    size_t scalar_size() const { return fp32() ? 4 : 8; }
    // SIPG on perturbed (non-axis-aligned) cells: the advection operator
    // kernel (without its diffusion terms) plus the general-geometry SIPG
    // kernel, then the inverse mass
    bool sipg_general() const
    {
      return cfg.operator_kind == DG_OP_ADVECTION_DIFFUSION &&
             cfg.perturbation != 0.;
    }
    void  *solver_vector(size_t idx, size_t bytes);
  };

  // kernels launched from host code
  // advection_only: the advection variant also for an advection-diffusion
  // operator (its diffusion terms come from aux_sipg_general)
  void launch_operator(const Operator &op,
                       const OpParams &P,
                       int             nc,
                       bool            fp32,
                       bool            skew,
                       int             fin,
                       int             epi,
                       const void     *io,
                       bool            advection_only = false);

  // the tile shape for a launch of variant `base_key` (shape bits 0):
  // forced, default, or tuned on first use per device and mesh size class
  TileChoice tile_choice(Operator &op, uint32_t base_key, const OpParams &P,
                         bool fp32, int epi, const void *io);
  // the tuner's time of every shape (-1: not built) for the operator's
  // mesh size class; false if the variant was not tuned
  bool tile_candidates(const Operator &op, uint32_t base_key,
                       double (&ms)[3]);
  int  tile_cells(int dim, int n, int shape);

  double velocity_time_factor(const dg_config &cfg, double t);

  // one stage of a low-storage Runge-Kutta step on three rotating
  // registers (0 = solution): stage time t + c dt, factors b dt and a dt,
  // reads registers r_in (stage input) and s_in (solution), writes s_out
  // and, unless last, r_out
  struct LsrkStage
  {
    double c, b, a;
    int    r_in, s_in, s_out, r_out;
    bool   last;
  };
  std::vector<LsrkStage> lsrk_plan(int scheme);
} // namespace dgops
