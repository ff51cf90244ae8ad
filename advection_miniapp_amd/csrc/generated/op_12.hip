#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 7, 4, 1, 0, 1, 0, 1, 0), &launch_op<3, 7, 4, float, 0, true, false, 1, 0>);
  static FdmRegistrar f1(fdm_key(3, 7, 3, 0), &launch_fdm<3, 7, 3, double, double, double>);
  static OpRegistrar r2(op_key(3, 7, 1, 1, 2, 0, 1, 0, 1), &launch_op<3, 7, 1, float, 2, false, true, 0, 1>);
  static FdmRegistrar f3(fdm_key(3, 6, 3, 1), &launch_fdm<3, 6, 3, float, float, float>);
  static OpRegistrar r4(op_key(3, 5, 4, 1, 2, 1, 0, 1, 0), &launch_op<3, 5, 4, float, 2, true, false, 1, 0>);
  static OpRegistrar r5(op_key(3, 5, 3, 1, 2, 1, 0, 1, 0), &launch_op<3, 5, 3, float, 2, true, false, 1, 0>);
  static FdmRegistrar f6(fdm_key(3, 6, 1, 0), &launch_fdm<3, 6, 1, double, double, double>);
  static FdmRegistrar f7(fdm_key(3, 5, 2, 0), &launch_fdm<3, 5, 2, double, double, double>);
  static OpRegistrar r8(op_key(2, 8, 1, 1, 2, 0, 1, 0, 1), &launch_op<2, 8, 1, float, 2, false, true, 0, 1>);
  static FdmRegistrar f9(fdm_key(3, 4, 3, 1), &launch_fdm<3, 4, 3, float, float, float>);
  static OpRegistrar r10(op_key(2, 6, 3, 1, 2, 1, 0, 1, 0), &launch_op<2, 6, 3, float, 2, true, false, 1, 0>);
  static OpRegistrar r11(op_key(2, 8, 1, 0, 0, 1, 0, 1, 0), &launch_op<2, 8, 1, double, 0, true, false, 1, 0>);
  static FdmRegistrar f12(fdm_key(3, 4, 2, 1), &launch_fdm<3, 4, 2, float, float, float>);
  static OpRegistrar r13(op_key(2, 6, 1, 1, 2, 1, 1, 1, 0), &launch_op<2, 6, 1, float, 2, true, true, 1, 0>);
  static OpRegistrar r14(op_key(3, 3, 4, 0, 0, 1, 0, 1, 0), &launch_op<3, 3, 4, double, 0, true, false, 1, 0>);
  static FdmRegistrar f15(fdm_key(3, 4, 1, 0), &launch_fdm<3, 4, 1, double, double, double>);
  static OpRegistrar r16(op_key(3, 3, 3, 0, 2, 1, 0, 1, 0), &launch_op<3, 3, 3, double, 2, true, false, 1, 0>);
  static OpRegistrar r17(op_key(2, 4, 3, 0, 2, 1, 0, 1, 0), &launch_op<2, 4, 3, double, 2, true, false, 1, 0>);
  static FdmRegistrar f18(fdm_key(3, 3, 2, 2), &launch_fdm<3, 3, 2, float, double, double>);
  static OpRegistrar r19(op_key(2, 5, 1, 1, 0, 0, 0, 0, 0), &launch_op<2, 5, 1, float, 0, false, false, 0, 0>);
  static OpRegistrar r20(op_key(3, 3, 1, 0, 2, 1, 0, 1, 0), &launch_op<3, 3, 1, double, 2, true, false, 1, 0>);
  static OpRegistrar r21(op_key(2, 4, 1, 1, 0, 1, 0, 1, 0), &launch_op<2, 4, 1, float, 0, true, false, 1, 0>);
  static OpRegistrar r22(op_key(2, 3, 1, 1, 1, 0, 1, 0, 1), &launch_op<2, 3, 1, float, 1, false, true, 0, 1>);
  static OpRegistrar r23(op_key(3, 2, 1, 0, 1, 0, 1, 0, 1), &launch_op<3, 2, 1, double, 1, false, true, 0, 1>);
  static OpRegistrar r24(op_key(2, 3, 1, 1, 0, 1, 0, 1, 0), &launch_op<2, 3, 1, float, 0, true, false, 1, 0>);
  static OpRegistrar r25(op_key(3, 2, 1, 1, 0, 1, 0, 1, 0), &launch_op<3, 2, 1, float, 0, true, false, 1, 0>);
  static FdmRegistrar f26(fdm_key(3, 2, 1, 2), &launch_fdm<3, 2, 1, float, double, double>);
} // namespace dgops
