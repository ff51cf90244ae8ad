#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 7, 4, 0), &launch_fdm<3, 7, 4, double, double, double>);
  static FdmRegistrar f1(fdm_key(3, 7, 3, 2), &launch_fdm<3, 7, 3, float, double, double>);
  static FdmRegistrar f2(fdm_key(3, 7, 2, 0), &launch_fdm<3, 7, 2, double, double, double>);
  static OpRegistrar r3(op_key(3, 6, 2, 0, 0, 1, 0, 1, 0), &launch_op<3, 6, 2, double, 0, true, false, 1, 0>);
  static FdmRegistrar f4(fdm_key(3, 5, 4, 1), &launch_fdm<3, 5, 4, float, float, float>);
  static FdmRegistrar f5(fdm_key(3, 5, 3, 1), &launch_fdm<3, 5, 3, float, float, float>);
  static FdmRegistrar f6(fdm_key(3, 6, 1, 2), &launch_fdm<3, 6, 1, float, double, double>);
  static FdmRegistrar f7(fdm_key(3, 5, 2, 2), &launch_fdm<3, 5, 2, float, double, double>);
  static FdmRegistrar f8(fdm_key(2, 8, 2, 0), &launch_fdm<2, 8, 2, double, double, double>);
  static OpRegistrar r9(op_key(2, 7, 2, 0, 0, 1, 0, 1, 0), &launch_op<2, 7, 2, double, 0, true, false, 1, 0>);
  static FdmRegistrar f10(fdm_key(2, 6, 3, 1), &launch_fdm<2, 6, 3, float, float, float>);
  static OpRegistrar r11(op_key(2, 8, 1, 0, 2, 1, 0, 1, 0), &launch_op<2, 8, 1, double, 2, true, false, 1, 0>);
  static OpRegistrar r12(op_key(2, 5, 4, 0, 0, 1, 0, 1, 0), &launch_op<2, 5, 4, double, 0, true, false, 1, 0>);
  static FdmRegistrar f13(fdm_key(2, 6, 2, 1), &launch_fdm<2, 6, 2, float, float, float>);
  static OpRegistrar r14(op_key(3, 3, 4, 1, 0, 1, 0, 1, 0), &launch_op<3, 3, 4, float, 0, true, false, 1, 0>);
  static FdmRegistrar f15(fdm_key(3, 4, 1, 2), &launch_fdm<3, 4, 1, float, double, double>);
  static OpRegistrar r16(op_key(3, 3, 3, 1, 2, 1, 0, 1, 0), &launch_op<3, 3, 3, float, 2, true, false, 1, 0>);
  static OpRegistrar r17(op_key(2, 4, 3, 1, 2, 1, 0, 1, 0), &launch_op<2, 4, 3, float, 2, true, false, 1, 0>);
  static OpRegistrar r18(op_key(2, 4, 2, 0, 2, 1, 0, 1, 0), &launch_op<2, 4, 2, double, 2, true, false, 1, 0>);
  static OpRegistrar r19(op_key(2, 5, 1, 0, 0, 1, 0, 1, 0), &launch_op<2, 5, 1, double, 0, true, false, 1, 0>);
  static OpRegistrar r20(op_key(2, 3, 3, 0, 0, 1, 0, 1, 0), &launch_op<2, 3, 3, double, 0, true, false, 1, 0>);
  static OpRegistrar r21(op_key(2, 4, 1, 1, 2, 1, 0, 1, 0), &launch_op<2, 4, 1, float, 2, true, false, 1, 0>);
  static OpRegistrar r22(op_key(2, 3, 1, 1, 2, 0, 1, 0, 0), &launch_op<2, 3, 1, float, 2, false, true, 0, 0>);
  static OpRegistrar r23(op_key(3, 2, 1, 0, 2, 0, 1, 0, 0), &launch_op<3, 2, 1, double, 2, false, true, 0, 0>);
  static OpRegistrar r24(op_key(2, 3, 1, 1, 2, 1, 0, 1, 0), &launch_op<2, 3, 1, float, 2, true, false, 1, 0>);
  static OpRegistrar r25(op_key(3, 2, 1, 1, 2, 1, 0, 1, 0), &launch_op<3, 2, 1, float, 2, true, false, 1, 0>);
  static OpRegistrar r26(op_key(2, 2, 1, 0, 0, 0, 0, 0, 0), &launch_op<2, 2, 1, double, 0, false, false, 0, 0>);
} // namespace dgops
