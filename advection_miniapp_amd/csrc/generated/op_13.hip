#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 7, 4, 1, 1, 1, 0, 1, 0), &launch_op<3, 7, 4, float, true, true, false, 1, 0>);
  static FdmRegistrar f1(fdm_key(3, 6, 4, 0), &launch_fdm<3, 6, 4, double, double, double>);
  static OpRegistrar r2(op_key(3, 8, 1, 0, 1, 1, 0, 1, 0), &launch_op<3, 8, 1, double, true, true, false, 1, 0>);
  static FdmRegistrar f3(fdm_key(3, 6, 2, 2), &launch_fdm<3, 6, 2, float, double, double>);
  static FdmRegistrar f4(fdm_key(2, 8, 4, 2), &launch_fdm<2, 8, 4, float, double, double>);
  static OpRegistrar r5(op_key(3, 6, 1, 0, 0, 1, 0, 1, 0), &launch_op<3, 6, 1, double, false, true, false, 1, 0>);
  static FdmRegistrar f6(fdm_key(3, 5, 2, 2), &launch_fdm<3, 5, 2, float, double, double>);
  static BurgersRegistrar g7(burgers_key(2, 5, 0, 0), &launch_burgers<2, 5, 10, false, 0>);
  static OpRegistrar r8(op_key(2, 7, 1, 0, 1, 1, 1, 1, 0), &launch_op<2, 7, 1, double, true, true, true, 1, 0>);
  static FdmRegistrar f9(fdm_key(3, 5, 1, 0), &launch_fdm<3, 5, 1, double, double, double>);
  static FdmRegistrar f10(fdm_key(2, 8, 1, 0), &launch_fdm<2, 8, 1, double, double, double>);
  static FdmRegistrar f11(fdm_key(2, 6, 2, 1), &launch_fdm<2, 6, 2, float, float, float>);
  static OpRegistrar r12(op_key(3, 3, 4, 1, 0, 1, 0, 1, 0), &launch_op<3, 3, 4, float, false, true, false, 1, 0>);
  static FdmRegistrar f13(fdm_key(3, 4, 1, 2), &launch_fdm<3, 4, 1, float, double, double>);
  static OpRegistrar r14(op_key(2, 6, 1, 0, 1, 0, 0, 0, 0), &launch_op<2, 6, 1, double, true, false, false, 0, 0>);
  static OpRegistrar r15(op_key(3, 3, 2, 1, 0, 1, 0, 1, 0), &launch_op<3, 3, 2, float, false, true, false, 1, 0>);
  static OpRegistrar r16(op_key(2, 5, 1, 0, 0, 0, 0, 0, 0), &launch_op<2, 5, 1, double, false, false, false, 0, 0>);
  static OpRegistrar r17(op_key(3, 3, 1, 1, 0, 0, 0, 0, 0), &launch_op<3, 3, 1, float, false, false, false, 0, 0>);
  static OpRegistrar r18(op_key(2, 4, 1, 0, 1, 0, 0, 0, 1), &launch_op<2, 4, 1, double, true, false, false, 0, 1>);
  static OpRegistrar r19(op_key(2, 3, 1, 1, 1, 0, 1, 0, 0), &launch_op<2, 3, 1, float, true, false, true, 0, 0>);
  static OpRegistrar r20(op_key(3, 2, 1, 1, 1, 0, 1, 0, 0), &launch_op<3, 2, 1, float, true, false, true, 0, 0>);
  static BurgersRegistrar g21(burgers_key(2, 3, 1, 0), &launch_burgers<2, 3, 3, true, 0>);
  static OpRegistrar r22(op_key(2, 2, 1, 0, 1, 0, 1, 0, 1), &launch_op<2, 2, 1, double, true, false, true, 0, 1>);
  static BurgersRegistrar g23(burgers_key(2, 2, 0, 0), &launch_burgers<2, 2, 1, false, 0>);
} // namespace dgops
