#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 7, 4, 1, 0, 1, 0, 1, 0), &launch_op<3, 7, 4, float, false, true, false, 1, 0>);
  static OpRegistrar r1(op_key(3, 6, 4, 1, 1, 1, 0, 1, 0), &launch_op<3, 6, 4, float, true, true, false, 1, 0>);
  static OpRegistrar r2(op_key(3, 8, 1, 1, 0, 1, 0, 1, 0), &launch_op<3, 8, 1, float, false, true, false, 1, 0>);
  static FdmRegistrar f3(fdm_key(3, 6, 2, 1), &launch_fdm<3, 6, 2, float, float, float>);
  static FdmRegistrar f4(fdm_key(2, 8, 4, 1), &launch_fdm<2, 8, 4, float, float, float>);
  static OpRegistrar r5(op_key(3, 6, 1, 1, 1, 0, 0, 0, 0), &launch_op<3, 6, 1, float, true, false, false, 0, 0>);
  static FdmRegistrar f6(fdm_key(3, 5, 2, 1), &launch_fdm<3, 5, 2, float, float, float>);
  static FdmRegistrar f7(fdm_key(3, 4, 4, 2), &launch_fdm<3, 4, 4, float, double, double>);
  static OpRegistrar r8(op_key(2, 7, 1, 0, 1, 0, 1, 0, 1), &launch_op<2, 7, 1, double, true, false, true, 0, 1>);
  static OpRegistrar r9(op_key(3, 5, 1, 1, 1, 1, 0, 1, 0), &launch_op<3, 5, 1, float, true, true, false, 1, 0>);
  static OpRegistrar r10(op_key(3, 4, 1, 1, 1, 1, 1, 1, 0), &launch_op<3, 4, 1, float, true, true, true, 1, 0>);
  static FdmRegistrar f11(fdm_key(2, 6, 2, 0), &launch_fdm<2, 6, 2, double, double, double>);
  static OpRegistrar r12(op_key(3, 3, 4, 0, 1, 1, 0, 1, 0), &launch_op<3, 3, 4, double, true, true, false, 1, 0>);
  static FdmRegistrar f13(fdm_key(3, 4, 1, 1), &launch_fdm<3, 4, 1, float, float, float>);
  static OpRegistrar r14(op_key(2, 6, 1, 0, 0, 0, 0, 0, 1), &launch_op<2, 6, 1, double, false, false, false, 0, 1>);
  static OpRegistrar r15(op_key(3, 3, 2, 0, 1, 1, 0, 1, 0), &launch_op<3, 3, 2, double, true, true, false, 1, 0>);
  static FdmRegistrar f16(fdm_key(2, 4, 2, 2), &launch_fdm<2, 4, 2, float, double, double>);
  static OpRegistrar r17(op_key(3, 3, 1, 0, 1, 0, 0, 0, 1), &launch_op<3, 3, 1, double, true, false, false, 0, 1>);
  static OpRegistrar r18(op_key(2, 4, 1, 0, 1, 0, 0, 0, 0), &launch_op<2, 4, 1, double, true, false, false, 0, 0>);
  static OpRegistrar r19(op_key(2, 3, 1, 0, 1, 1, 1, 1, 0), &launch_op<2, 3, 1, double, true, true, true, 1, 0>);
  static OpRegistrar r20(op_key(3, 2, 1, 0, 1, 1, 1, 1, 0), &launch_op<3, 2, 1, double, true, true, true, 1, 0>);
  static BurgersRegistrar g21(burgers_key(2, 3, 0, 1), &launch_burgers<2, 3, 3, false, 1>);
  static OpRegistrar r22(op_key(2, 2, 1, 0, 1, 0, 1, 0, 0), &launch_op<2, 2, 1, double, true, false, true, 0, 0>);
  static FdmRegistrar f23(fdm_key(2, 2, 1, 2), &launch_fdm<2, 2, 1, float, double, double>);
} // namespace dgops
