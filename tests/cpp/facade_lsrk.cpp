// Drop-in check of include/dgops.hpp: a caller-side, templated low-storage
// Runge-Kutta integrator written against the reference's operator concept
// (perform_stage with the aliasing of advection_solver.cc:906-939) drives
// the GPU operator through the C++ facade. Writes the solution after
// `steps` steps of LSRK(5,4) to the file named by argv[1] (raw doubles).
// argv[3] picks the integrator: "caller" (default, the one below),
// "library" (dgops::LowStorageRungeKuttaIntegrator, one call per step) or
// "steps" (its perform_time_steps, all steps in one call).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "dgops.hpp"

namespace
{
  // the caller's integrator: stage 0 reads and updates `solution` and
  // writes vec_ri; later stages read vec_ri and overwrite it
  class LowStorageRungeKutta
  {
  public:
    explicit LowStorageRungeKutta(int scheme)
    {
      const int n = dg_lsrk_n_stages(scheme);
      bi.resize(n);
      ai.resize(n);
      dgops::check(dg_lsrk_coefficients(scheme, bi.data(), ai.data()) == n ?
                     DG_OK :
                     DG_ERR_ARGUMENT);
    }

    template <typename V, typename Operator>
    void
    perform_time_step(const Operator &op, double t, double dt, V &solution,
                      V &vec_ri, V &vec_ki) const
    {
      const std::size_t s = bi.size();
      op.perform_stage(t, bi[0] * dt, ai[0] * dt, solution, vec_ri, solution,
                       vec_ri);
      double sum_b = 0;
      for (std::size_t i = 1; i < s; ++i)
        {
          const double ci = sum_b + ai[i - 1];
          op.perform_stage(t + ci * dt, bi[i] * dt,
                           i + 1 == s ? 0. : ai[i] * dt, vec_ri, vec_ki,
                           solution, vec_ri);
          sum_b += bi[i - 1];
        }
    }

  private:
    std::vector<double> bi, ai;
  };
} // namespace

int
main(int argc, char **argv)
{
  if (argc < 2)
    {
      std::fprintf(stderr,
                   "usage: %s output.bin [steps] [caller|library|steps]\n",
                   argv[0]);
      return 2;
    }
  const int         steps = argc > 2 ? std::atoi(argv[2]) : 3;
  const std::string mode  = argc > 3 ? argv[3] : "caller";
  try
    {
      dg_config cfg;
      dg_config_default(&cfg);
      cfg.dim              = 3;
      cfg.degree           = 3;
      cfg.refine           = 2;
      cfg.perturbation     = 0.1;
      cfg.velocity_kind    = DG_VEL_CELLULAR_3D;
      cfg.solution_kind    = DG_SOL_GAUSSIAN;
      cfg.solution_param[0] = 0.5;
      cfg.solution_param[1] = 0.5;
      cfg.solution_param[2] = 0.5;
      cfg.solution_param[3] = 0.15;
      dgops::AdvectionOperation<double> op(cfg);
      dgops::DeviceVector<double>       u, ri, ki;
      op.initialize_dof_vector(u);
      op.initialize_dof_vector(ri);
      op.initialize_dof_vector(ki);
      op.project_initial(u, 0.);
      const double dt = 2e-3;
      if (mode == "caller")
        {
          LowStorageRungeKutta lsrk(DG_LSRK_5_4);
          for (int s = 0; s < steps; ++s)
            lsrk.perform_time_step(op, s * dt, dt, u, ri, ki);
        }
      else if (mode == "library")
        {
          dgops::LowStorageRungeKuttaIntegrator lsrk(DG_LSRK_5_4);
          for (int s = 0; s < steps; ++s)
            lsrk.perform_time_step(op, s * dt, dt, u, ri, ki);
        }
      else if (mode == "steps")
        {
          dgops::LowStorageRungeKuttaIntegrator lsrk(DG_LSRK_5_4);
          lsrk.perform_time_steps(op, 0., dt, steps, u, ri, ki);
        }
      else
        throw std::runtime_error("unknown mode " + mode);
      std::vector<double> host(u.size());
      if (hipMemcpy(host.data(), u.data(), sizeof(double) * u.size(),
                    hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("copy back failed");
      std::FILE *f = std::fopen(argv[1], "wb");
      if (!f)
        throw std::runtime_error("cannot open output");
      std::fwrite(host.data(), sizeof(double), host.size(), f);
      std::fclose(f);
      const auto me = op.compute_mass_and_energy(u);
      std::printf("facade_lsrk: %zu dofs, %d steps, mass %.6e energy %.6e\n",
                  u.size(), steps, me[0], me[1]);
    }
  catch (const std::exception &e)
    {
      std::fprintf(stderr, "facade_lsrk: %s\n", e.what());
      return 1;
    }
  return 0;
}
