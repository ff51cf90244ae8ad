#include "../registry.hpp"

namespace dgops
{
  static BurgersRegistrar g0(burgers_key(3, 5, 1, 0), &launch_burgers<3, 5, 10, true, 0>);
  static FdmRegistrar f1(fdm_key(3, 7, 3, 2), &launch_fdm<3, 7, 3, float, double, double>);
  static OpRegistrar r2(op_key(3, 8, 1, 0, 1, 0, 0, 0, 1), &launch_op<3, 8, 1, double, true, false, false, 0, 1>);
  static OpRegistrar r3(op_key(3, 6, 1, 1, 1, 0, 1, 0, 0), &launch_op<3, 6, 1, float, true, false, true, 0, 0>);
  static FdmRegistrar f4(fdm_key(3, 7, 1, 1), &launch_fdm<3, 7, 1, float, float, float>);
  static OpRegistrar r5(op_key(3, 6, 1, 0, 0, 0, 0, 0, 1), &launch_op<3, 6, 1, double, false, false, false, 0, 1>);
  static OpRegistrar r6(op_key(3, 5, 1, 1, 1, 0, 1, 0, 0), &launch_op<3, 5, 1, float, true, false, true, 0, 0>);
  static FdmRegistrar f7(fdm_key(2, 8, 2, 1), &launch_fdm<2, 8, 2, float, float, float>);
  static OpRegistrar r8(op_key(2, 7, 2, 0, 1, 1, 0, 1, 0), &launch_op<2, 7, 2, double, true, true, false, 1, 0>);
  static OpRegistrar r9(op_key(3, 5, 1, 1, 1, 0, 0, 0, 0), &launch_op<3, 5, 1, float, true, false, false, 0, 0>);
  static OpRegistrar r10(op_key(3, 4, 1, 0, 1, 0, 1, 0, 1), &launch_op<3, 4, 1, double, true, false, true, 0, 1>);
  static OpRegistrar r11(op_key(2, 6, 1, 0, 1, 1, 1, 1, 0), &launch_op<2, 6, 1, double, true, true, true, 1, 0>);
  static FdmRegistrar f12(fdm_key(2, 7, 1, 0), &launch_fdm<2, 7, 1, double, double, double>);
  static FdmRegistrar f13(fdm_key(2, 4, 4, 0), &launch_fdm<2, 4, 4, double, double, double>);
  static FdmRegistrar f14(fdm_key(3, 3, 3, 0), &launch_fdm<3, 3, 3, double, double, double>);
  static FdmRegistrar f15(fdm_key(2, 4, 3, 0), &launch_fdm<2, 4, 3, double, double, double>);
  static OpRegistrar r16(op_key(2, 4, 1, 1, 1, 0, 1, 0, 1), &launch_op<2, 4, 1, float, true, false, true, 0, 1>);
  static FdmRegistrar f17(fdm_key(2, 3, 4, 2), &launch_fdm<2, 3, 4, float, double, double>);
  static BurgersRegistrar g18(burgers_key(3, 3, 1, 1), &launch_burgers<3, 3, 3, true, 1>);
  static OpRegistrar r19(op_key(2, 3, 2, 1, 0, 1, 0, 1, 0), &launch_op<2, 3, 2, float, false, true, false, 1, 0>);
  static OpRegistrar r20(op_key(3, 2, 2, 1, 0, 1, 0, 1, 0), &launch_op<3, 2, 2, float, false, true, false, 1, 0>);
  static FdmRegistrar f21(fdm_key(2, 3, 1, 0), &launch_fdm<2, 3, 1, double, double, double>);
  static OpRegistrar r22(op_key(2, 2, 2, 0, 0, 1, 0, 1, 0), &launch_op<2, 2, 2, double, false, true, false, 1, 0>);
  static OpRegistrar r23(op_key(2, 2, 1, 0, 1, 1, 0, 1, 0), &launch_op<2, 2, 1, double, true, true, false, 1, 0>);
} // namespace dgops
