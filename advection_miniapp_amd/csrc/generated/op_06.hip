#include "../registry.hpp"

namespace dgops
{
  static BurgersRegistrar g0(burgers_key(3, 5, 0, 0), &launch_burgers<3, 5, 10, false, 0>);
  static FdmRegistrar f1(fdm_key(3, 7, 3, 0), &launch_fdm<3, 7, 3, double, double, double>);
  static OpRegistrar r2(op_key(3, 8, 1, 0, 0, 0, 0, 0, 1), &launch_op<3, 8, 1, double, false, false, false, 0, 1>);
  static OpRegistrar r3(op_key(3, 6, 1, 0, 1, 0, 1, 0, 1), &launch_op<3, 6, 1, double, true, false, true, 0, 1>);
  static OpRegistrar r4(op_key(3, 7, 1, 1, 1, 1, 0, 1, 0), &launch_op<3, 7, 1, float, true, true, false, 1, 0>);
  static FdmRegistrar f5(fdm_key(2, 7, 4, 2), &launch_fdm<2, 7, 4, float, double, double>);
  static OpRegistrar r6(op_key(3, 5, 1, 0, 1, 0, 1, 0, 1), &launch_op<3, 5, 1, double, true, false, true, 0, 1>);
  static OpRegistrar r7(op_key(2, 8, 1, 1, 1, 1, 1, 1, 0), &launch_op<2, 8, 1, float, true, true, true, 1, 0>);
  static FdmRegistrar f8(fdm_key(3, 4, 3, 2), &launch_fdm<3, 4, 3, float, double, double>);
  static OpRegistrar r9(op_key(3, 5, 1, 0, 1, 0, 0, 0, 1), &launch_op<3, 5, 1, double, true, false, false, 0, 1>);
  static OpRegistrar r10(op_key(3, 4, 2, 1, 1, 1, 0, 1, 0), &launch_op<3, 4, 2, float, true, true, false, 1, 0>);
  static OpRegistrar r11(op_key(2, 6, 1, 0, 1, 0, 1, 0, 0), &launch_op<2, 6, 1, double, true, false, true, 0, 0>);
  static OpRegistrar r12(op_key(2, 7, 1, 0, 1, 1, 0, 1, 0), &launch_op<2, 7, 1, double, true, true, false, 1, 0>);
  static OpRegistrar r13(op_key(2, 4, 4, 1, 0, 1, 0, 1, 0), &launch_op<2, 4, 4, float, false, true, false, 1, 0>);
  static OpRegistrar r14(op_key(3, 3, 3, 1, 0, 1, 0, 1, 0), &launch_op<3, 3, 3, float, false, true, false, 1, 0>);
  static OpRegistrar r15(op_key(2, 4, 3, 1, 0, 1, 0, 1, 0), &launch_op<2, 4, 3, float, false, true, false, 1, 0>);
  static OpRegistrar r16(op_key(2, 4, 1, 0, 1, 1, 1, 1, 0), &launch_op<2, 4, 1, double, true, true, true, 1, 0>);
  static FdmRegistrar f17(fdm_key(2, 3, 4, 0), &launch_fdm<2, 3, 4, double, double, double>);
  static BurgersRegistrar g18(burgers_key(3, 3, 0, 1), &launch_burgers<3, 3, 3, false, 1>);
  static OpRegistrar r19(op_key(2, 3, 2, 0, 0, 1, 0, 1, 0), &launch_op<2, 3, 2, double, false, true, false, 1, 0>);
  static OpRegistrar r20(op_key(3, 2, 2, 0, 0, 1, 0, 1, 0), &launch_op<3, 2, 2, double, false, true, false, 1, 0>);
  static OpRegistrar r21(op_key(2, 3, 1, 0, 1, 1, 0, 1, 0), &launch_op<2, 3, 1, double, true, true, false, 1, 0>);
  static OpRegistrar r22(op_key(3, 2, 1, 0, 1, 1, 0, 1, 0), &launch_op<3, 2, 1, double, true, true, false, 1, 0>);
  static OpRegistrar r23(op_key(2, 2, 1, 0, 0, 1, 0, 1, 0), &launch_op<2, 2, 1, double, false, true, false, 1, 0>);
} // namespace dgops
