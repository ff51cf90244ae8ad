/*
 * C++ facade over the dgops C ABI with the member functions of the
 * reference's operator classes, so that templated integrators and solver
 * drivers written against AdvectionOperation compile unchanged:
 *
 *   apply(src, dst, t)                       advection_solver.cc:374-377
 *   perform_stage(t, b_dt, a_dt, ri, ki,
 *                 solution, next_ri)          advection_solver.cc:379-386
 *   project_initial(dst)                     advection_solver.cc:760-784
 *   compute_mass_and_energy(vec)             advection_solver.cc:787-818
 *   set_time(t, dt), vmult(dst, src),
 *   apply_inverse_mass_matrix(dst),
 *   precondition_block_jacobi(dst, src)      advection_solver_implicit_real.cc:869-909
 *   compute_rhs, vmult, update_solution,
 *   precondition_block_jacobi (stage blocks)  advection_solver_irk_real.cc:1266-1331
 *
 * Vectors: any type for which dgops::device_data(v) returns the device
 * pointer of its contiguous storage (overloads below cover
 * dgops::DeviceVector and every type with a data() member; a deal.II
 * LinearAlgebra::distributed::Vector<Number, MemorySpace::Default> user
 * adds one overload returning v.get_values()). Errors throw
 * std::runtime_error with the library's message, as the reference's
 * AssertThrow does.
 */
#ifndef DGOPS_HPP
#define DGOPS_HPP

#include <hip/hip_runtime.h>

#include <array>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <utility>

#include "dgops.h"

namespace dgops
{
  inline void
  check(int rc)
  {
    if (rc != DG_OK)
      throw std::runtime_error(std::string("dgops: ") + dg_last_error());
  }

  // owning device vector, the stand-in for the reference's
  // LinearAlgebra::distributed::Vector on one GPU
  template <typename Number>
  class DeviceVector
  {
  public:
    DeviceVector() = default;
    explicit DeviceVector(std::size_t n) { reinit(n); }
    DeviceVector(const DeviceVector &o) { *this = o; }
    DeviceVector(DeviceVector &&o) noexcept
      : ptr_(std::exchange(o.ptr_, nullptr))
      , n_(std::exchange(o.n_, 0))
    {}
    DeviceVector &
    operator=(const DeviceVector &o)
    {
      if (this != &o)
        {
          reinit(o.n_);
          if (n_ > 0 &&
              hipMemcpy(ptr_, o.ptr_, n_ * sizeof(Number),
                        hipMemcpyDeviceToDevice) != hipSuccess)
            throw std::runtime_error("dgops: device copy failed");
        }
      return *this;
    }
    ~DeviceVector() { release(); }

    void
    reinit(std::size_t n)
    {
      if (n == n_ && ptr_)
        return;
      release();
      n_ = n;
      if (n_ > 0 &&
          hipMalloc(reinterpret_cast<void **>(&ptr_), n_ * sizeof(Number)) !=
            hipSuccess)
        throw std::runtime_error("dgops: hipMalloc failed");
      if (n_ > 0 && hipMemset(ptr_, 0, n_ * sizeof(Number)) != hipSuccess)
        throw std::runtime_error("dgops: hipMemset failed");
    }
    std::size_t size() const { return n_; }
    Number *data() { return ptr_; }
    const Number *data() const { return ptr_; }

  private:
    void
    release()
    {
      if (ptr_)
        (void)hipFree(ptr_);
      ptr_ = nullptr;
      n_   = 0;
    }
    Number     *ptr_ = nullptr;
    std::size_t n_   = 0;
  };

  template <typename V>
  inline auto
  device_data(V &v) -> decltype(v.data())
  {
    return v.data();
  }

  template <typename V>
  inline auto
  device_data(const V &v) -> decltype(v.data())
  {
    return v.data();
  }

  // one GPU operator: explicit, implicit and IRK interfaces of the
  // reference's AdvectionOperation (which of them apply depends on the
  // configuration, as in the reference's separate programs)
  template <typename Number>
  class AdvectionOperation
  {
  public:
    explicit AdvectionOperation(const dg_config &cfg,
                                hipStream_t      stream = nullptr)
    {
      check(dg_create(&cfg, stream, &op_));
      check(dg_get_info(op_, &info_));
      stages_ = cfg.irk_stages > 0 ? cfg.irk_stages : 1;
    }
    AdvectionOperation(const AdvectionOperation &) = delete;
    AdvectionOperation &operator=(const AdvectionOperation &) = delete;
    ~AdvectionOperation()
    {
      if (op_)
        dg_destroy(op_);
    }

    const dg_info &info() const { return info_; }
    dg_operator *handle() const { return op_; }

    template <typename V>
    void
    initialize_dof_vector(V &v) const
    {
      v.reinit(info_.vector_size);
    }

    template <typename V>
    void
    initialize_block_vector(V &v) const
    {
      v.reinit(info_.vector_size * stages_);
    }

    // ---- explicit ----------------------------------------------------------
    template <typename V>
    void
    apply(const V &src, V &dst, const double current_time)
    {
      check(dg_apply(op_, device_data(src), device_data(dst), current_time));
    }

    template <typename V>
    void
    perform_stage(const Number current_time,
                  const Number factor_solution,
                  const Number factor_ai,
                  const V     &current_ri,
                  V           &vec_ki,
                  V           &solution,
                  V           &next_ri) const
    {
      check(dg_perform_stage(op_, current_time, factor_solution, factor_ai,
                             device_data(current_ri), device_data(vec_ki),
                             device_data(solution), device_data(next_ri)));
    }

    template <typename V>
    void
    project_initial(V &dst, const double time = 0.) const
    {
      check(dg_project_initial(op_, device_data(dst), time));
    }

    template <typename V>
    std::array<double, 3>
    compute_mass_and_energy(const V &vec) const
    {
      std::array<double, 3> out;
      check(dg_compute_mass_and_energy(op_, device_data(vec), out.data()));
      return out;
    }

    template <typename V>
    std::array<double, 2>
    l2_error(const V &vec, const double time) const
    {
      std::array<double, 2> out;
      check(dg_l2_error(op_, device_data(vec), time, out.data()));
      return out;
    }

    // ---- implicit single stage ---------------------------------------------
    void
    set_time(const double current_time, const double time_step) const
    {
      check(dg_set_time(op_, current_time, time_step));
    }

    template <typename V>
    void
    vmult(V &dst, const V &src) const
    {
      if (dst.size() == info_.vector_size * (std::size_t)stages_ &&
          stages_ > 1)
        check(dg_irk_vmult(op_, device_data(dst), device_data(src)));
      else
        check(dg_vmult(op_, device_data(dst), device_data(src)));
    }

    template <typename V>
    void
    apply_inverse_mass_matrix(V &dst) const
    {
      check(dg_apply_inverse_mass_matrix(op_, device_data(dst)));
    }

    // the reference's do_batched_solver switch
    // (advection_solver_irk_real.cc:124, :2108-2126): the block-Jacobi step
    // on stage blocks runs one GMRES step per cell in the collocation basis
    // instead of the FDM alone
    void set_batched_solver(const bool on) { batched_ = on; }

    template <typename V>
    void
    precondition_block_jacobi(V &dst, const V &src) const
    {
      if (dst.size() == info_.vector_size * (std::size_t)stages_ &&
          stages_ > 1)
        check(batched_ ? dg_irk_precondition_cell_solver(op_, device_data(dst),
                                                         device_data(src)) :
                         dg_irk_precondition_block_jacobi(op_,
                                                          device_data(dst),
                                                          device_data(src)));
      else
        check(dg_precondition_block_jacobi(op_, device_data(dst),
                                           device_data(src)));
    }

    // ---- implicit Runge-Kutta ------------------------------------------------
    template <typename BV, typename V>
    void
    compute_rhs(BV &dst, const V &src) const
    {
      check(dg_irk_compute_rhs(op_, device_data(dst), device_data(src)));
    }

    template <typename V, typename BV>
    void
    update_solution(V &solution, const BV &stage_solution) const
    {
      check(dg_irk_update_solution(op_, device_data(solution),
                                   device_data(stage_solution)));
    }

  private:
    dg_operator *op_ = nullptr;
    dg_info      info_{};
    int          stages_  = 1;
    bool         batched_ = false;
  };

  // the reference's LowStorageRungeKuttaIntegrator (advection_solver.cc
  // :823-944) with the stages fused inside the library: perform_time_step
  // is one dg_lsrk_time_step; perform_time_steps runs the driver's time
  // loop (advection_solver.cc:1264-1299) in one call, replaying a captured
  // HIP graph of a step when the operator does not depend on time
  class LowStorageRungeKuttaIntegrator
  {
  public:
    explicit LowStorageRungeKuttaIntegrator(const int scheme)
      : scheme_(scheme)
    {
      if (dg_lsrk_n_stages(scheme) < 0)
        check(DG_ERR_ARGUMENT);
    }

    unsigned int n_stages() const { return dg_lsrk_n_stages(scheme_); }

    template <typename Op, typename V>
    void
    perform_time_step(const Op &op, const double current_time,
                      const double time_step, V &solution, V &vec_ri,
                      V &vec_ki) const
    {
      check(dg_lsrk_time_step(op.handle(), scheme_, current_time, time_step,
                              device_data(solution), device_data(vec_ri),
                              device_data(vec_ki)));
    }

    template <typename Op, typename V>
    void
    perform_time_steps(const Op &op, const double current_time,
                       const double time_step, const int n_steps,
                       V &solution, V &vec_ri, V &vec_ki) const
    {
      check(dg_lsrk_time_steps(op.handle(), scheme_, current_time, time_step,
                               n_steps, device_data(solution),
                               device_data(vec_ri), device_data(vec_ki)));
    }

  private:
    int scheme_;
  };

  // BlockJacobi wrapper of the reference (advection_solver_implicit_real.cc
  // :1540-1556): vmult(dst, src) forwards to precondition_block_jacobi
  template <typename Operator>
  class BlockJacobi
  {
  public:
    explicit BlockJacobi(const Operator &op)
      : op_(op)
    {}
    template <typename V>
    void
    vmult(V &dst, const V &src) const
    {
      op_.precondition_block_jacobi(dst, src);
    }

  private:
    const Operator &op_;
  };
} // namespace dgops

#endif
