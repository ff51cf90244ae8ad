#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 2, 1), &launch_fdm<3, 8, 2, float, float, float>);
  static OpRegistrar r1(op_key(3, 8, 1, 0, 0, 0, 0, 0, 1), &launch_op<3, 8, 1, double, false, false, false, 0, 1>);
  static FdmRegistrar f2(fdm_key(3, 6, 2, 0), &launch_fdm<3, 6, 2, double, double, double>);
  static FdmRegistrar f3(fdm_key(2, 8, 4, 2), &launch_fdm<2, 8, 4, float, double, double>);
  static OpRegistrar r4(op_key(3, 6, 1, 1, 1, 1, 0, 1, 0), &launch_op<3, 6, 1, float, true, true, false, 1, 0>);
  static FdmRegistrar f5(fdm_key(2, 7, 3, 1), &launch_fdm<2, 7, 3, float, float, float>);
  static FdmRegistrar f6(fdm_key(2, 6, 4, 2), &launch_fdm<2, 6, 4, float, double, double>);
  static OpRegistrar r7(op_key(3, 5, 1, 0, 1, 0, 0, 0, 0), &launch_op<3, 5, 1, double, true, false, false, 0, 0>);
  static OpRegistrar r8(op_key(3, 4, 1, 1, 1, 0, 1, 0, 0), &launch_op<3, 4, 1, float, true, false, true, 0, 0>);
  static FdmRegistrar f9(fdm_key(2, 6, 2, 2), &launch_fdm<2, 6, 2, float, double, double>);
  static OpRegistrar r10(op_key(3, 4, 1, 0, 0, 0, 0, 0, 1), &launch_op<3, 4, 1, double, false, false, false, 0, 1>);
  static FdmRegistrar f11(fdm_key(2, 5, 2, 0), &launch_fdm<2, 5, 2, double, double, double>);
  static OpRegistrar r12(op_key(2, 4, 3, 0, 1, 1, 0, 1, 0), &launch_op<2, 4, 3, double, true, true, false, 1, 0>);
  static FdmRegistrar f13(fdm_key(2, 4, 2, 0), &launch_fdm<2, 4, 2, double, double, double>);
  static OpRegistrar r14(op_key(3, 3, 1, 1, 1, 0, 0, 0, 0), &launch_op<3, 3, 1, float, true, false, false, 0, 0>);
  static OpRegistrar r15(op_key(2, 4, 1, 0, 1, 1, 0, 1, 0), &launch_op<2, 4, 1, double, true, true, false, 1, 0>);
  static FdmRegistrar f16(fdm_key(3, 2, 3, 1), &launch_fdm<3, 2, 3, float, float, float>);
  static OpRegistrar r17(op_key(2, 3, 1, 1, 0, 1, 0, 1, 0), &launch_op<2, 3, 1, float, false, true, false, 1, 0>);
  static OpRegistrar r18(op_key(2, 2, 2, 0, 1, 1, 0, 1, 0), &launch_op<2, 2, 2, double, true, true, false, 1, 0>);
  static FdmRegistrar f19(fdm_key(2, 2, 1, 1), &launch_fdm<2, 2, 1, float, float, float>);
} // namespace dgops
