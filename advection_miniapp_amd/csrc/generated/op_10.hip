#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 7, 4, 0, 0, 1, 0, 1, 0), &launch_op<3, 7, 4, double, false, true, false, 1, 0>);
  static OpRegistrar r1(op_key(3, 6, 4, 0, 1, 1, 0, 1, 0), &launch_op<3, 6, 4, double, true, true, false, 1, 0>);
  static OpRegistrar r2(op_key(3, 8, 1, 1, 1, 0, 0, 0, 0), &launch_op<3, 8, 1, float, true, false, false, 0, 0>);
  static OpRegistrar r3(op_key(3, 6, 1, 1, 1, 1, 1, 1, 0), &launch_op<3, 6, 1, float, true, true, true, 1, 0>);
  static BatchRegistrar s4(batch_key(3, 7), &launch_batch<3, 7>);
  static OpRegistrar r5(op_key(3, 6, 1, 0, 1, 0, 0, 0, 1), &launch_op<3, 6, 1, double, true, false, false, 0, 1>);
  static OpRegistrar r6(op_key(3, 5, 1, 1, 1, 1, 1, 1, 0), &launch_op<3, 5, 1, float, true, true, true, 1, 0>);
  static FdmRegistrar f7(fdm_key(3, 4, 4, 0), &launch_fdm<3, 4, 4, double, double, double>);
  static OpRegistrar r8(op_key(2, 7, 2, 1, 1, 1, 0, 1, 0), &launch_op<2, 7, 2, float, true, true, false, 1, 0>);
  static OpRegistrar r9(op_key(3, 5, 1, 1, 0, 1, 0, 1, 0), &launch_op<3, 5, 1, float, false, true, false, 1, 0>);
  static OpRegistrar r10(op_key(3, 4, 1, 1, 1, 0, 1, 0, 0), &launch_op<3, 4, 1, float, true, false, true, 0, 0>);
  static OpRegistrar r11(op_key(2, 6, 1, 1, 1, 0, 1, 0, 1), &launch_op<2, 6, 1, float, true, false, true, 0, 1>);
  static FdmRegistrar f12(fdm_key(2, 7, 1, 2), &launch_fdm<2, 7, 1, float, double, double>);
  static FdmRegistrar f13(fdm_key(2, 4, 4, 2), &launch_fdm<2, 4, 4, float, double, double>);
  static FdmRegistrar f14(fdm_key(3, 3, 3, 2), &launch_fdm<3, 3, 3, float, double, double>);
  static FdmRegistrar f15(fdm_key(2, 4, 3, 2), &launch_fdm<2, 4, 3, float, double, double>);
  static FdmRegistrar f16(fdm_key(2, 4, 2, 0), &launch_fdm<2, 4, 2, double, double, double>);
  static OpRegistrar r17(op_key(3, 3, 1, 0, 0, 0, 0, 0, 1), &launch_op<3, 3, 1, double, false, false, false, 0, 1>);
  static OpRegistrar r18(op_key(2, 4, 1, 0, 0, 0, 0, 0, 0), &launch_op<2, 4, 1, double, false, false, false, 0, 0>);
  static OpRegistrar r19(op_key(2, 3, 1, 0, 1, 0, 1, 0, 0), &launch_op<2, 3, 1, double, true, false, true, 0, 0>);
  static OpRegistrar r20(op_key(3, 2, 1, 0, 1, 0, 1, 0, 0), &launch_op<3, 2, 1, double, true, false, true, 0, 0>);
  static FdmRegistrar f21(fdm_key(2, 3, 1, 2), &launch_fdm<2, 3, 1, float, double, double>);
  static OpRegistrar r22(op_key(2, 2, 2, 1, 0, 1, 0, 1, 0), &launch_op<2, 2, 2, float, false, true, false, 1, 0>);
  static FdmRegistrar f23(fdm_key(2, 2, 1, 0), &launch_fdm<2, 2, 1, double, double, double>);
} // namespace dgops
