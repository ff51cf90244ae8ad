#include "../registry.hpp"

namespace dgops
{
  static BurgersRegistrar g0(burgers_key(3, 5, 1, 1), &launch_burgers<3, 5, 10, true, 1>);
  static OpRegistrar r1(op_key(3, 6, 4, 0, 0, 1, 0, 1, 0), &launch_op<3, 6, 4, double, false, true, false, 1, 0>);
  static OpRegistrar r2(op_key(3, 8, 1, 1, 1, 0, 0, 0, 0), &launch_op<3, 8, 1, float, true, false, false, 0, 0>);
  static OpRegistrar r3(op_key(3, 6, 1, 1, 1, 1, 1, 1, 0), &launch_op<3, 6, 1, float, true, true, true, 1, 0>);
  static FdmRegistrar f4(fdm_key(2, 8, 4, 1), &launch_fdm<2, 8, 4, float, float, float>);
  static OpRegistrar r5(op_key(3, 6, 1, 0, 0, 1, 0, 1, 0), &launch_op<3, 6, 1, double, false, true, false, 1, 0>);
  static OpRegistrar r6(op_key(2, 7, 3, 0, 0, 1, 0, 1, 0), &launch_op<2, 7, 3, double, false, true, false, 1, 0>);
  static BurgersRegistrar g7(burgers_key(2, 5, 0, 1), &launch_burgers<2, 5, 10, false, 1>);
  static OpRegistrar r8(op_key(2, 7, 1, 1, 1, 0, 1, 0, 0), &launch_op<2, 7, 1, float, true, false, true, 0, 0>);
  static FdmRegistrar f9(fdm_key(3, 5, 1, 2), &launch_fdm<3, 5, 1, float, double, double>);
  static FdmRegistrar f10(fdm_key(3, 4, 2, 1), &launch_fdm<3, 4, 2, float, float, float>);
  static OpRegistrar r11(op_key(2, 5, 3, 1, 0, 1, 0, 1, 0), &launch_op<2, 5, 3, float, false, true, false, 1, 0>);
  static OpRegistrar r12(op_key(3, 4, 1, 0, 0, 0, 0, 0, 0), &launch_op<3, 4, 1, double, false, false, false, 0, 0>);
  static OpRegistrar r13(op_key(2, 5, 1, 0, 1, 0, 1, 0, 1), &launch_op<2, 5, 1, double, true, false, true, 0, 1>);
  static FdmRegistrar f14(fdm_key(2, 6, 1, 0), &launch_fdm<2, 6, 1, double, double, double>);
  static FdmRegistrar f15(fdm_key(3, 3, 2, 0), &launch_fdm<3, 3, 2, double, double, double>);
  static OpRegistrar r16(op_key(2, 5, 1, 1, 1, 1, 0, 1, 0), &launch_op<2, 5, 1, float, true, true, false, 1, 0>);
  static FdmRegistrar f17(fdm_key(2, 3, 3, 0), &launch_fdm<2, 3, 3, double, double, double>);
  static FdmRegistrar f18(fdm_key(2, 4, 1, 1), &launch_fdm<2, 4, 1, float, float, float>);
  static FdmRegistrar f19(fdm_key(3, 2, 3, 2), &launch_fdm<3, 2, 3, float, double, double>);
  static OpRegistrar r20(op_key(2, 3, 1, 0, 1, 0, 0, 0, 1), &launch_op<2, 3, 1, double, true, false, false, 0, 1>);
  static OpRegistrar r21(op_key(3, 2, 1, 1, 1, 0, 0, 0, 0), &launch_op<3, 2, 1, float, true, false, false, 0, 0>);
  static OpRegistrar r22(op_key(2, 2, 1, 0, 1, 0, 0, 0, 1), &launch_op<2, 2, 1, double, true, false, false, 0, 1>);
} // namespace dgops
