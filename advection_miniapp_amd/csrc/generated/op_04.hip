#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 3, 1), &launch_fdm<3, 8, 3, float, float, float>);
  static OpRegistrar r1(op_key(3, 7, 3, 1, 0, 1, 0, 1, 0), &launch_op<3, 7, 3, float, false, true, false, 1, 0>);
  static FdmRegistrar f2(fdm_key(3, 7, 2, 2), &launch_fdm<3, 7, 2, float, double, double>);
  static OpRegistrar r3(op_key(3, 6, 2, 1, 1, 1, 0, 1, 0), &launch_op<3, 6, 2, float, true, true, false, 1, 0>);
  static OpRegistrar r4(op_key(3, 7, 1, 1, 0, 1, 0, 1, 0), &launch_op<3, 7, 1, float, false, true, false, 1, 0>);
  static FdmRegistrar f5(fdm_key(2, 7, 4, 0), &launch_fdm<2, 7, 4, double, double, double>);
  static OpRegistrar r6(op_key(3, 5, 2, 1, 1, 1, 0, 1, 0), &launch_op<3, 5, 2, float, true, true, false, 1, 0>);
  static OpRegistrar r7(op_key(2, 8, 1, 1, 1, 0, 1, 0, 0), &launch_op<2, 8, 1, float, true, false, true, 0, 0>);
  static FdmRegistrar f8(fdm_key(3, 4, 3, 0), &launch_fdm<3, 4, 3, double, double, double>);
  static OpRegistrar r9(op_key(3, 5, 1, 0, 0, 0, 0, 0, 1), &launch_op<3, 5, 1, double, false, false, false, 0, 1>);
  static OpRegistrar r10(op_key(3, 4, 2, 0, 1, 1, 0, 1, 0), &launch_op<3, 4, 2, double, true, true, false, 1, 0>);
  static OpRegistrar r11(op_key(2, 6, 2, 1, 0, 1, 0, 1, 0), &launch_op<2, 6, 2, float, false, true, false, 1, 0>);
  static OpRegistrar r12(op_key(2, 7, 1, 0, 0, 1, 0, 1, 0), &launch_op<2, 7, 1, double, false, true, false, 1, 0>);
  static OpRegistrar r13(op_key(2, 4, 4, 0, 0, 1, 0, 1, 0), &launch_op<2, 4, 4, double, false, true, false, 1, 0>);
  static OpRegistrar r14(op_key(3, 3, 3, 0, 0, 1, 0, 1, 0), &launch_op<3, 3, 3, double, false, true, false, 1, 0>);
  static OpRegistrar r15(op_key(2, 4, 3, 0, 0, 1, 0, 1, 0), &launch_op<2, 4, 3, double, false, true, false, 1, 0>);
  static OpRegistrar r16(op_key(2, 4, 1, 0, 1, 0, 1, 0, 0), &launch_op<2, 4, 1, double, true, false, true, 0, 0>);
  static OpRegistrar r17(op_key(2, 3, 4, 1, 0, 1, 0, 1, 0), &launch_op<2, 3, 4, float, false, true, false, 1, 0>);
  static FdmRegistrar f18(fdm_key(3, 3, 1, 2), &launch_fdm<3, 3, 1, float, double, double>);
  static FdmRegistrar f19(fdm_key(3, 2, 4, 1), &launch_fdm<3, 2, 4, float, float, float>);
  static OpRegistrar r20(op_key(2, 2, 4, 1, 0, 1, 0, 1, 0), &launch_op<2, 2, 4, float, false, true, false, 1, 0>);
  static OpRegistrar r21(op_key(2, 3, 1, 0, 0, 1, 0, 1, 0), &launch_op<2, 3, 1, double, false, true, false, 1, 0>);
  static OpRegistrar r22(op_key(3, 2, 1, 0, 0, 1, 0, 1, 0), &launch_op<3, 2, 1, double, false, true, false, 1, 0>);
  static OpRegistrar r23(op_key(2, 2, 1, 1, 0, 0, 0, 0, 0), &launch_op<2, 2, 1, float, false, false, false, 0, 0>);
} // namespace dgops
