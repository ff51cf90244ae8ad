// Drop-in check of include/dgops.hpp on the IRK side of the reference's
// operator concept (advection_solver_irk_real.cc:1266-1331): stage-block
// vmult, compute_rhs, update_solution and the BlockJacobi preconditioner,
// with and without the do_batched_solver switch, called the way a deal.II
// solver calls them (vmult(dst, src) on the operator and on the
// preconditioner). Writes, as raw doubles to argv[1]: vmult(z), the
// block-Jacobi step of z with the FDM alone and with the batched cell
// solver, compute_rhs(u) and update_solution(u, z).
#include <cstdio>
#include <vector>

#include "dgops.hpp"

namespace
{
  template <typename V>
  void
  append(std::vector<double> &out, const V &v)
  {
    const std::size_t n0 = out.size();
    out.resize(n0 + v.size());
    if (hipMemcpy(out.data() + n0, v.data(), sizeof(double) * v.size(),
                  hipMemcpyDeviceToHost) != hipSuccess)
      throw std::runtime_error("copy back failed");
  }
} // namespace

int
main(int argc, char **argv)
{
  if (argc < 2)
    {
      std::fprintf(stderr, "usage: %s output.bin\n", argv[0]);
      return 2;
    }
  try
    {
      dg_config cfg;
      dg_config_default(&cfg);
      cfg.dim               = 3;
      cfg.degree            = 3;
      cfg.refine            = 2;
      cfg.perturbation      = 0.1;
      cfg.velocity_kind     = DG_VEL_CELLULAR_3D;
      cfg.factor_skew       = 0.5;
      cfg.irk_stages        = 3;
      cfg.solution_kind     = DG_SOL_GAUSSIAN;
      cfg.solution_param[0] = 0.5;
      cfg.solution_param[1] = 0.5;
      cfg.solution_param[2] = 0.5;
      cfg.solution_param[3] = 0.15;
      dgops::AdvectionOperation<double> op(cfg);
      dgops::DeviceVector<double>       u, z, dst;
      op.initialize_dof_vector(u);
      op.initialize_block_vector(z);
      op.initialize_block_vector(dst);
      op.project_initial(u, 0.);
      op.set_time(0.1, 0.05);
      // a stage-block vector with data in every block
      op.compute_rhs(z, u);
      std::vector<double> out;
      op.vmult(dst, z);
      append(out, dst);
      const dgops::BlockJacobi<dgops::AdvectionOperation<double>> prec(op);
      prec.vmult(dst, z);
      append(out, dst);
      op.set_batched_solver(true);
      prec.vmult(dst, z);
      append(out, dst);
      op.set_batched_solver(false);
      op.compute_rhs(dst, u);
      append(out, dst);
      op.update_solution(u, z);
      append(out, u);
      std::FILE *f = std::fopen(argv[1], "wb");
      if (!f)
        throw std::runtime_error("cannot open output");
      std::fwrite(out.data(), sizeof(double), out.size(), f);
      std::fclose(f);
      std::printf("facade_irk: %zu unknowns per stage block\n", z.size());
    }
  catch (const std::exception &e)
    {
      std::fprintf(stderr, "facade_irk: %s\n", e.what());
      return 1;
    }
  return 0;
}
