#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 3, 0), &launch_fdm<3, 8, 3, double, double, double>);
  static OpRegistrar r1(op_key(3, 7, 3, 0, 1, 1, 0, 1, 0), &launch_op<3, 7, 3, double, true, true, false, 1, 0>);
  static FdmRegistrar f2(fdm_key(3, 7, 2, 1), &launch_fdm<3, 7, 2, float, float, float>);
  static OpRegistrar r3(op_key(3, 6, 2, 1, 0, 1, 0, 1, 0), &launch_op<3, 6, 2, float, false, true, false, 1, 0>);
  static OpRegistrar r4(op_key(3, 7, 1, 0, 0, 1, 0, 1, 0), &launch_op<3, 7, 1, double, false, true, false, 1, 0>);
  static OpRegistrar r5(op_key(2, 7, 4, 1, 1, 1, 0, 1, 0), &launch_op<2, 7, 4, float, true, true, false, 1, 0>);
  static OpRegistrar r6(op_key(3, 5, 2, 1, 0, 1, 0, 1, 0), &launch_op<3, 5, 2, float, false, true, false, 1, 0>);
  static OpRegistrar r7(op_key(2, 8, 1, 0, 1, 1, 1, 1, 0), &launch_op<2, 8, 1, double, true, true, true, 1, 0>);
  static OpRegistrar r8(op_key(3, 4, 3, 1, 1, 1, 0, 1, 0), &launch_op<3, 4, 3, float, true, true, false, 1, 0>);
  static OpRegistrar r9(op_key(3, 5, 1, 0, 0, 0, 0, 0, 0), &launch_op<3, 5, 1, double, false, false, false, 0, 0>);
  static OpRegistrar r10(op_key(3, 4, 2, 0, 0, 1, 0, 1, 0), &launch_op<3, 4, 2, double, false, true, false, 1, 0>);
  static OpRegistrar r11(op_key(2, 6, 2, 0, 1, 1, 0, 1, 0), &launch_op<2, 6, 2, double, true, true, false, 1, 0>);
  static OpRegistrar r12(op_key(2, 7, 1, 1, 1, 0, 0, 0, 0), &launch_op<2, 7, 1, float, true, false, false, 0, 0>);
  static OpRegistrar r13(op_key(3, 4, 1, 1, 1, 1, 0, 1, 0), &launch_op<3, 4, 1, float, true, true, false, 1, 0>);
  static FdmRegistrar f14(fdm_key(2, 5, 2, 2), &launch_fdm<2, 5, 2, float, double, double>);
  static BurgersRegistrar g15(burgers_key(2, 4, 1, 1), &launch_burgers<2, 4, 6, true, 1>);
  static OpRegistrar r16(op_key(2, 4, 2, 1, 1, 1, 0, 1, 0), &launch_op<2, 4, 2, float, true, true, false, 1, 0>);
  static OpRegistrar r17(op_key(2, 3, 4, 0, 1, 1, 0, 1, 0), &launch_op<2, 3, 4, double, true, true, false, 1, 0>);
  static FdmRegistrar f18(fdm_key(3, 3, 1, 1), &launch_fdm<3, 3, 1, float, float, float>);
  static FdmRegistrar f19(fdm_key(3, 2, 4, 0), &launch_fdm<3, 2, 4, double, double, double>);
  static OpRegistrar r20(op_key(2, 2, 4, 0, 1, 1, 0, 1, 0), &launch_op<2, 2, 4, double, true, true, false, 1, 0>);
  static OpRegistrar r21(op_key(2, 3, 1, 1, 1, 0, 0, 0, 0), &launch_op<2, 3, 1, float, true, false, false, 0, 0>);
  static OpRegistrar r22(op_key(3, 2, 1, 1, 1, 0, 0, 0, 0), &launch_op<3, 2, 1, float, true, false, false, 0, 0>);
  static OpRegistrar r23(op_key(2, 2, 1, 0, 1, 0, 0, 0, 1), &launch_op<2, 2, 1, double, true, false, false, 0, 1>);
} // namespace dgops
