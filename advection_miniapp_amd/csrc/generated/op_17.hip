#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 8, 1, 0, 1, 0, 1, 0, 0), &launch_op<3, 8, 1, double, 1, false, true, 0, 0>);
  static OpRegistrar r1(op_key(3, 6, 4, 1, 0, 1, 0, 1, 0), &launch_op<3, 6, 4, float, 0, true, false, 1, 0>);
  static OpRegistrar r2(op_key(3, 8, 1, 0, 0, 0, 0, 0, 0), &launch_op<3, 8, 1, double, 0, false, false, 0, 0>);
  static OpRegistrar r3(op_key(3, 6, 2, 0, 2, 1, 0, 1, 0), &launch_op<3, 6, 2, double, 2, true, false, 1, 0>);
  static FdmRegistrar f4(fdm_key(3, 5, 4, 2), &launch_fdm<3, 5, 4, float, double, double>);
  static FdmRegistrar f5(fdm_key(3, 5, 3, 0), &launch_fdm<3, 5, 3, double, double, double>);
  static FdmRegistrar f6(fdm_key(3, 6, 1, 1), &launch_fdm<3, 6, 1, float, float, float>);
  static OpRegistrar r7(op_key(3, 5, 1, 1, 2, 1, 1, 1, 0), &launch_op<3, 5, 1, float, 2, true, true, 1, 0>);
  static OpRegistrar r8(op_key(2, 8, 1, 1, 2, 0, 1, 0, 0), &launch_op<2, 8, 1, float, 2, false, true, 0, 0>);
  static FdmRegistrar f9(fdm_key(3, 4, 3, 0), &launch_fdm<3, 4, 3, double, double, double>);
  static OpRegistrar r10(op_key(2, 6, 3, 1, 0, 1, 0, 1, 0), &launch_op<2, 6, 3, float, 0, true, false, 1, 0>);
  static OpRegistrar r11(op_key(2, 8, 1, 0, 2, 0, 0, 0, 1), &launch_op<2, 8, 1, double, 2, false, false, 0, 1>);
  static FdmRegistrar f12(fdm_key(2, 8, 1, 1), &launch_fdm<2, 8, 1, float, float, float>);
  static OpRegistrar r13(op_key(2, 6, 1, 1, 1, 0, 1, 0, 0), &launch_op<2, 6, 1, float, 1, false, true, 0, 0>);
  static OpRegistrar r14(op_key(2, 7, 1, 0, 2, 1, 0, 1, 0), &launch_op<2, 7, 1, double, 2, true, false, 1, 0>);
  static OpRegistrar r15(op_key(2, 4, 4, 0, 0, 1, 0, 1, 0), &launch_op<2, 4, 4, double, 0, true, false, 1, 0>);
  static OpRegistrar r16(op_key(2, 5, 1, 1, 1, 0, 1, 0, 1), &launch_op<2, 5, 1, float, 1, false, true, 0, 1>);
  static OpRegistrar r17(op_key(2, 6, 1, 1, 2, 1, 0, 1, 0), &launch_op<2, 6, 1, float, 2, true, false, 1, 0>);
  static OpRegistrar r18(op_key(3, 3, 1, 0, 2, 0, 1, 0, 0), &launch_op<3, 3, 1, double, 2, false, true, 0, 0>);
  static OpRegistrar r19(op_key(2, 4, 1, 1, 1, 1, 1, 1, 0), &launch_op<2, 4, 1, float, 1, true, true, 1, 0>);
  static OpRegistrar r20(op_key(2, 3, 4, 1, 0, 1, 0, 1, 0), &launch_op<2, 3, 4, float, 0, true, false, 1, 0>);
  static FdmRegistrar f21(fdm_key(3, 3, 1, 2), &launch_fdm<3, 3, 1, float, double, double>);
  static FdmRegistrar f22(fdm_key(2, 4, 1, 2), &launch_fdm<2, 4, 1, float, double, double>);
  static FdmRegistrar f23(fdm_key(2, 3, 2, 2), &launch_fdm<2, 3, 2, float, double, double>);
  static OpRegistrar r24(op_key(3, 2, 1, 1, 1, 1, 1, 1, 0), &launch_op<3, 2, 1, float, 1, true, true, 1, 0>);
  static BurgersRegistrar g25(burgers_key(2, 3, 0, 1), &launch_burgers<2, 3, 3, false, 1>);
  static OpRegistrar r26(op_key(2, 2, 2, 1, 0, 1, 0, 1, 0), &launch_op<2, 2, 2, float, 0, true, false, 1, 0>);
  static OpRegistrar r27(op_key(2, 2, 1, 0, 0, 0, 0, 0, 1), &launch_op<2, 2, 1, double, 0, false, false, 0, 1>);
} // namespace dgops
