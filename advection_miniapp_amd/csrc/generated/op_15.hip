#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 7, 4, 1), &launch_fdm<3, 7, 4, float, float, float>);
  static FdmRegistrar f1(fdm_key(3, 6, 4, 2), &launch_fdm<3, 6, 4, float, double, double>);
  static FdmRegistrar f2(fdm_key(3, 8, 1, 0), &launch_fdm<3, 8, 1, double, double, double>);
  static OpRegistrar r3(op_key(3, 5, 4, 0, 1, 1, 0, 1, 0), &launch_op<3, 5, 4, double, true, true, false, 1, 0>);
  static OpRegistrar r4(op_key(3, 5, 3, 0, 1, 1, 0, 1, 0), &launch_op<3, 5, 3, double, true, true, false, 1, 0>);
  static OpRegistrar r5(op_key(3, 6, 1, 0, 1, 1, 0, 1, 0), &launch_op<3, 6, 1, double, true, true, false, 1, 0>);
  static OpRegistrar r6(op_key(2, 7, 3, 0, 1, 1, 0, 1, 0), &launch_op<2, 7, 3, double, true, true, false, 1, 0>);
  static BurgersRegistrar g7(burgers_key(2, 5, 1, 0), &launch_burgers<2, 5, 10, true, 0>);
  static OpRegistrar r8(op_key(2, 7, 1, 1, 1, 0, 1, 0, 1), &launch_op<2, 7, 1, float, true, false, true, 0, 1>);
  static FdmRegistrar f9(fdm_key(3, 5, 1, 2), &launch_fdm<3, 5, 1, float, double, double>);
  static FdmRegistrar f10(fdm_key(2, 8, 1, 2), &launch_fdm<2, 8, 1, float, double, double>);
  static OpRegistrar r11(op_key(2, 5, 3, 0, 0, 1, 0, 1, 0), &launch_op<2, 5, 3, double, false, true, false, 1, 0>);
  static FdmRegistrar f12(fdm_key(3, 3, 4, 0), &launch_fdm<3, 3, 4, double, double, double>);
  static OpRegistrar r13(op_key(2, 5, 2, 0, 0, 1, 0, 1, 0), &launch_op<2, 5, 2, double, false, true, false, 1, 0>);
  static OpRegistrar r14(op_key(2, 6, 1, 1, 0, 0, 0, 0, 0), &launch_op<2, 6, 1, float, false, false, false, 0, 0>);
  static OpRegistrar r15(op_key(3, 3, 1, 0, 1, 0, 1, 0, 0), &launch_op<3, 3, 1, double, true, false, true, 0, 0>);
  static OpRegistrar r16(op_key(2, 5, 1, 0, 1, 0, 0, 0, 0), &launch_op<2, 5, 1, double, true, false, false, 0, 0>);
  static OpRegistrar r17(op_key(3, 3, 1, 0, 0, 1, 0, 1, 0), &launch_op<3, 3, 1, double, false, true, false, 1, 0>);
  static OpRegistrar r18(op_key(2, 4, 1, 1, 1, 0, 0, 0, 0), &launch_op<2, 4, 1, float, true, false, false, 0, 0>);
  static OpRegistrar r19(op_key(2, 3, 1, 1, 1, 1, 1, 1, 0), &launch_op<2, 3, 1, float, true, true, true, 1, 0>);
  static OpRegistrar r20(op_key(3, 2, 1, 1, 1, 1, 1, 1, 0), &launch_op<3, 2, 1, float, true, true, true, 1, 0>);
  static OpRegistrar r21(op_key(2, 2, 3, 0, 0, 1, 0, 1, 0), &launch_op<2, 2, 3, double, false, true, false, 1, 0>);
  static OpRegistrar r22(op_key(2, 2, 1, 1, 1, 0, 1, 0, 0), &launch_op<2, 2, 1, float, true, false, true, 0, 0>);
  static BurgersRegistrar g23(burgers_key(2, 2, 1, 0), &launch_burgers<2, 2, 1, true, 0>);
} // namespace dgops
