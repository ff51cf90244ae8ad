#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 7, 3, 0, 1, 1, 1, 0), &launch_op<3, 7, 3, double, true, true, 1, 0>);
  static OpRegistrar r1(op_key(3, 8, 1, 0, 0, 0, 0, 1), &launch_op<3, 8, 1, double, false, false, 0, 1>);
  static OpRegistrar r2(op_key(3, 6, 2, 1, 1, 1, 1, 0), &launch_op<3, 6, 2, float, true, true, 1, 0>);
  static OpRegistrar r3(op_key(3, 7, 1, 1, 0, 1, 1, 0), &launch_op<3, 7, 1, float, false, true, 1, 0>);
  static FdmRegistrar f4(fdm_key(2, 8, 3, 0), &launch_fdm<2, 8, 3, double, double, double>);
  static FdmRegistrar f5(fdm_key(3, 6, 1, 0), &launch_fdm<3, 6, 1, double, double, double>);
  static OpRegistrar r6(op_key(3, 4, 4, 1, 1, 1, 1, 0), &launch_op<3, 4, 4, float, true, true, 1, 0>);
  static FdmRegistrar f7(fdm_key(3, 4, 3, 2), &launch_fdm<3, 4, 3, float, double, double>);
  static OpRegistrar r8(op_key(3, 5, 1, 1, 0, 1, 1, 0), &launch_op<3, 5, 1, float, false, true, 1, 0>);
  static FdmRegistrar f9(fdm_key(2, 8, 1, 0), &launch_fdm<2, 8, 1, double, double, double>);
  static OpRegistrar r10(op_key(2, 5, 3, 0, 1, 1, 1, 0), &launch_op<2, 5, 3, double, true, true, 1, 0>);
  static OpRegistrar r11(op_key(3, 3, 4, 0, 1, 1, 1, 0), &launch_op<3, 3, 4, double, true, true, 1, 0>);
  static FdmRegistrar f12(fdm_key(2, 4, 4, 1), &launch_fdm<2, 4, 4, float, float, float>);
  static OpRegistrar r13(op_key(2, 6, 1, 0, 0, 0, 0, 1), &launch_op<2, 6, 1, double, false, false, 0, 1>);
  static OpRegistrar r14(op_key(3, 3, 2, 1, 1, 1, 1, 0), &launch_op<3, 3, 2, float, true, true, 1, 0>);
  static OpRegistrar r15(op_key(2, 5, 1, 1, 1, 1, 1, 0), &launch_op<2, 5, 1, float, true, true, 1, 0>);
  static OpRegistrar r16(op_key(3, 3, 1, 1, 1, 1, 1, 0), &launch_op<3, 3, 1, float, true, true, 1, 0>);
  static OpRegistrar r17(op_key(2, 4, 1, 1, 1, 1, 1, 0), &launch_op<2, 4, 1, float, true, true, 1, 0>);
  static FdmRegistrar f18(fdm_key(3, 2, 3, 0), &launch_fdm<3, 2, 3, double, double, double>);
  static OpRegistrar r19(op_key(2, 3, 1, 0, 1, 0, 0, 1), &launch_op<2, 3, 1, double, true, false, 0, 1>);
  static OpRegistrar r20(op_key(3, 2, 1, 0, 1, 0, 0, 1), &launch_op<3, 2, 1, double, true, false, 0, 1>);
  static OpRegistrar r21(op_key(2, 2, 1, 0, 1, 0, 0, 1), &launch_op<2, 2, 1, double, true, false, 0, 1>);
} // namespace dgops
