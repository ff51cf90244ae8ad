#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 7, 4, 1, 2, 1, 0, 1, 0), &launch_op<3, 7, 4, float, 2, true, false, 1, 0>);
  static FdmRegistrar f1(fdm_key(3, 7, 3, 1), &launch_fdm<3, 7, 3, float, float, float>);
  static OpRegistrar r2(op_key(3, 7, 1, 1, 2, 1, 1, 1, 0), &launch_op<3, 7, 1, float, 2, true, true, 1, 0>);
  static FdmRegistrar f3(fdm_key(3, 6, 3, 2), &launch_fdm<3, 6, 3, float, double, double>);
  static FdmRegistrar f4(fdm_key(3, 5, 4, 0), &launch_fdm<3, 5, 4, double, double, double>);
  static FdmRegistrar f5(fdm_key(3, 5, 3, 0), &launch_fdm<3, 5, 3, double, double, double>);
  static FdmRegistrar f6(fdm_key(3, 6, 1, 1), &launch_fdm<3, 6, 1, float, float, float>);
  static FdmRegistrar f7(fdm_key(3, 5, 2, 1), &launch_fdm<3, 5, 2, float, float, float>);
  static OpRegistrar r8(op_key(2, 8, 1, 1, 2, 1, 1, 1, 0), &launch_op<2, 8, 1, float, 2, true, true, 1, 0>);
  static FdmRegistrar f9(fdm_key(3, 4, 3, 2), &launch_fdm<3, 4, 3, float, double, double>);
  static FdmRegistrar f10(fdm_key(2, 6, 3, 0), &launch_fdm<2, 6, 3, double, double, double>);
  static OpRegistrar r11(op_key(2, 8, 1, 1, 0, 1, 0, 1, 0), &launch_op<2, 8, 1, float, 0, true, false, 1, 0>);
  static FdmRegistrar f12(fdm_key(3, 4, 2, 2), &launch_fdm<3, 4, 2, float, double, double>);
  static FdmRegistrar f13(fdm_key(2, 6, 2, 0), &launch_fdm<2, 6, 2, double, double, double>);
  static OpRegistrar r14(op_key(3, 3, 4, 0, 2, 1, 0, 1, 0), &launch_op<3, 3, 4, double, 2, true, false, 1, 0>);
  static FdmRegistrar f15(fdm_key(3, 4, 1, 1), &launch_fdm<3, 4, 1, float, float, float>);
  static OpRegistrar r16(op_key(3, 3, 3, 1, 0, 1, 0, 1, 0), &launch_op<3, 3, 3, float, 0, true, false, 1, 0>);
  static OpRegistrar r17(op_key(2, 4, 3, 1, 0, 1, 0, 1, 0), &launch_op<2, 4, 3, float, 0, true, false, 1, 0>);
  static OpRegistrar r18(op_key(2, 4, 2, 0, 0, 1, 0, 1, 0), &launch_op<2, 4, 2, double, 0, true, false, 1, 0>);
  static OpRegistrar r19(op_key(2, 5, 1, 1, 2, 0, 0, 0, 0), &launch_op<2, 5, 1, float, 2, false, false, 0, 0>);
  static OpRegistrar r20(op_key(3, 3, 1, 1, 2, 1, 0, 1, 0), &launch_op<3, 3, 1, float, 2, true, false, 1, 0>);
  static OpRegistrar r21(op_key(2, 4, 1, 0, 2, 1, 0, 1, 0), &launch_op<2, 4, 1, double, 2, true, false, 1, 0>);
  static OpRegistrar r22(op_key(2, 3, 1, 1, 1, 1, 1, 1, 0), &launch_op<2, 3, 1, float, 1, true, true, 1, 0>);
  static OpRegistrar r23(op_key(3, 2, 1, 0, 1, 1, 1, 1, 0), &launch_op<3, 2, 1, double, 1, true, true, 1, 0>);
  static OpRegistrar r24(op_key(2, 3, 1, 0, 2, 1, 0, 1, 0), &launch_op<2, 3, 1, double, 2, true, false, 1, 0>);
  static OpRegistrar r25(op_key(3, 2, 1, 0, 2, 1, 0, 1, 0), &launch_op<3, 2, 1, double, 2, true, false, 1, 0>);
  static BatchRegistrar s26(batch_key(3, 2), &launch_batch<3, 2>);
} // namespace dgops
