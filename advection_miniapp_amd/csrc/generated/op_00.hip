#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 4, 0), &launch_fdm<3, 8, 4, double, double, double>);
  static OpRegistrar r1(op_key(3, 7, 3, 0, 1, 1, 0, 1, 0), &launch_op<3, 7, 3, double, true, true, false, 1, 0>);
  static OpRegistrar r2(op_key(3, 8, 1, 0, 1, 0, 0, 0, 1), &launch_op<3, 8, 1, double, true, false, false, 0, 1>);
  static FdmRegistrar f3(fdm_key(3, 6, 2, 2), &launch_fdm<3, 6, 2, float, double, double>);
  static OpRegistrar r4(op_key(3, 5, 3, 0, 0, 1, 0, 1, 0), &launch_op<3, 5, 3, double, false, true, false, 1, 0>);
  static FdmRegistrar f5(fdm_key(3, 6, 1, 1), &launch_fdm<3, 6, 1, float, float, float>);
  static OpRegistrar r6(op_key(3, 4, 4, 0, 1, 1, 0, 1, 0), &launch_op<3, 4, 4, double, true, true, false, 1, 0>);
  static OpRegistrar r7(op_key(3, 4, 3, 1, 1, 1, 0, 1, 0), &launch_op<3, 4, 3, float, true, true, false, 1, 0>);
  static OpRegistrar r8(op_key(3, 5, 1, 1, 1, 0, 0, 0, 0), &launch_op<3, 5, 1, float, true, false, false, 0, 0>);
  static OpRegistrar r9(op_key(3, 4, 1, 1, 1, 1, 1, 1, 0), &launch_op<3, 4, 1, float, true, true, true, 1, 0>);
  static OpRegistrar r10(op_key(2, 5, 3, 1, 1, 1, 0, 1, 0), &launch_op<2, 5, 3, float, true, true, false, 1, 0>);
  static OpRegistrar r11(op_key(3, 4, 1, 0, 1, 0, 0, 0, 1), &launch_op<3, 4, 1, double, true, false, false, 0, 1>);
  static FdmRegistrar f12(fdm_key(2, 5, 2, 2), &launch_fdm<2, 5, 2, float, double, double>);
  static FdmRegistrar f13(fdm_key(2, 4, 3, 0), &launch_fdm<2, 4, 3, double, double, double>);
  static FdmRegistrar f14(fdm_key(2, 4, 2, 2), &launch_fdm<2, 4, 2, float, double, double>);
  static OpRegistrar r15(op_key(3, 3, 1, 1, 0, 1, 0, 1, 0), &launch_op<3, 3, 1, float, false, true, false, 1, 0>);
  static OpRegistrar r16(op_key(3, 2, 4, 0, 1, 1, 0, 1, 0), &launch_op<3, 2, 4, double, true, true, false, 1, 0>);
  static OpRegistrar r17(op_key(2, 2, 4, 0, 1, 1, 0, 1, 0), &launch_op<2, 2, 4, double, true, true, false, 1, 0>);
  static OpRegistrar r18(op_key(2, 3, 1, 1, 1, 1, 0, 1, 0), &launch_op<2, 3, 1, float, true, true, false, 1, 0>);
  static OpRegistrar r19(op_key(2, 2, 1, 0, 1, 0, 1, 0, 0), &launch_op<2, 2, 1, double, true, false, true, 0, 0>);
  static BurgersRegistrar g20(burgers_key(2, 2, 0, 0), &launch_burgers<2, 2, 1, false, 0>);
} // namespace dgops
