#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 7, 4, 2), &launch_fdm<3, 7, 4, float, double, double>);
  static OpRegistrar r1(op_key(3, 6, 4, 0, 2, 1, 0, 1, 0), &launch_op<3, 6, 4, double, 2, true, false, 1, 0>);
  static FdmRegistrar f2(fdm_key(3, 7, 2, 2), &launch_fdm<3, 7, 2, float, double, double>);
  static OpRegistrar r3(op_key(3, 6, 2, 1, 0, 1, 0, 1, 0), &launch_op<3, 6, 2, float, 0, true, false, 1, 0>);
  static OpRegistrar r4(op_key(3, 7, 1, 0, 0, 0, 0, 0, 0), &launch_op<3, 7, 1, double, 0, false, false, 0, 0>);
  static FdmRegistrar f5(fdm_key(2, 8, 3, 0), &launch_fdm<2, 8, 3, double, double, double>);
  static BurgersRegistrar g6(burgers_key(3, 4, 0, 1), &launch_burgers<3, 4, 6, false, 1>);
  static OpRegistrar r7(op_key(2, 7, 3, 0, 2, 1, 0, 1, 0), &launch_op<2, 7, 3, double, 2, true, false, 1, 0>);
  static FdmRegistrar f8(fdm_key(2, 8, 2, 2), &launch_fdm<2, 8, 2, float, double, double>);
  static OpRegistrar r9(op_key(2, 7, 2, 1, 0, 1, 0, 1, 0), &launch_op<2, 7, 2, float, 0, true, false, 1, 0>);
  static OpRegistrar r10(op_key(3, 5, 1, 0, 0, 0, 0, 0, 0), &launch_op<3, 5, 1, double, 0, false, false, 0, 0>);
  static OpRegistrar r11(op_key(3, 4, 2, 0, 0, 1, 0, 1, 0), &launch_op<3, 4, 2, double, 0, true, false, 1, 0>);
  static OpRegistrar r12(op_key(2, 5, 4, 1, 0, 1, 0, 1, 0), &launch_op<2, 5, 4, float, 0, true, false, 1, 0>);
  static OpRegistrar r13(op_key(2, 5, 3, 0, 0, 1, 0, 1, 0), &launch_op<2, 5, 3, double, 0, true, false, 1, 0>);
  static FdmRegistrar f14(fdm_key(3, 3, 4, 0), &launch_fdm<3, 3, 4, double, double, double>);
  static OpRegistrar r15(op_key(2, 5, 2, 0, 0, 1, 0, 1, 0), &launch_op<2, 5, 2, double, 0, true, false, 1, 0>);
  static FdmRegistrar f16(fdm_key(3, 3, 3, 1), &launch_fdm<3, 3, 3, float, float, float>);
  static FdmRegistrar f17(fdm_key(2, 4, 3, 1), &launch_fdm<2, 4, 3, float, float, float>);
  static OpRegistrar r18(op_key(2, 4, 2, 1, 2, 1, 0, 1, 0), &launch_op<2, 4, 2, float, 2, true, false, 1, 0>);
  static OpRegistrar r19(op_key(2, 5, 1, 0, 2, 1, 0, 1, 0), &launch_op<2, 5, 1, double, 2, true, false, 1, 0>);
  static OpRegistrar r20(op_key(2, 3, 3, 1, 0, 1, 0, 1, 0), &launch_op<2, 3, 3, float, 0, true, false, 1, 0>);
  static OpRegistrar r21(op_key(3, 2, 4, 0, 2, 1, 0, 1, 0), &launch_op<3, 2, 4, double, 2, true, false, 1, 0>);
  static OpRegistrar r22(op_key(2, 3, 1, 1, 2, 1, 1, 1, 0), &launch_op<2, 3, 1, float, 2, true, true, 1, 0>);
  static OpRegistrar r23(op_key(3, 2, 1, 0, 2, 1, 1, 1, 0), &launch_op<3, 2, 1, double, 2, true, true, 1, 0>);
  static FdmRegistrar f24(fdm_key(2, 3, 1, 1), &launch_fdm<2, 3, 1, float, float, float>);
  static OpRegistrar r25(op_key(2, 2, 2, 0, 2, 1, 0, 1, 0), &launch_op<2, 2, 2, double, 2, true, false, 1, 0>);
  static OpRegistrar r26(op_key(2, 2, 1, 0, 2, 0, 0, 0, 0), &launch_op<2, 2, 1, double, 2, false, false, 0, 0>);
} // namespace dgops
