#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 8, 1, 1, 1, 1, 1, 1, 0), &launch_op<3, 8, 1, float, true, true, true, 1, 0>);
  static OpRegistrar r1(op_key(3, 7, 1, 0, 1, 1, 1, 1, 0), &launch_op<3, 7, 1, double, true, true, true, 1, 0>);
  static FdmRegistrar f2(fdm_key(3, 6, 3, 0), &launch_fdm<3, 6, 3, double, double, double>);
  static OpRegistrar r3(op_key(3, 7, 1, 0, 0, 0, 0, 0, 1), &launch_op<3, 7, 1, double, false, false, false, 0, 1>);
  static FdmRegistrar f4(fdm_key(2, 8, 3, 1), &launch_fdm<2, 8, 3, float, float, float>);
  static BurgersRegistrar g5(burgers_key(3, 4, 1, 0), &launch_burgers<3, 4, 6, true, 0>);
  static OpRegistrar r6(op_key(3, 4, 4, 0, 1, 1, 0, 1, 0), &launch_op<3, 4, 4, double, true, true, false, 1, 0>);
  static FdmRegistrar f7(fdm_key(2, 6, 4, 1), &launch_fdm<2, 6, 4, float, float, float>);
  static OpRegistrar r8(op_key(2, 6, 3, 1, 0, 1, 0, 1, 0), &launch_op<2, 6, 3, float, false, true, false, 1, 0>);
  static OpRegistrar r9(op_key(2, 8, 1, 1, 1, 0, 0, 0, 0), &launch_op<2, 8, 1, float, true, false, false, 0, 0>);
  static OpRegistrar r10(op_key(2, 5, 4, 1, 1, 1, 0, 1, 0), &launch_op<2, 5, 4, float, true, true, false, 1, 0>);
  static OpRegistrar r11(op_key(2, 7, 1, 0, 0, 0, 0, 0, 0), &launch_op<2, 7, 1, double, false, false, false, 0, 0>);
  static OpRegistrar r12(op_key(3, 4, 1, 1, 0, 0, 0, 0, 0), &launch_op<3, 4, 1, float, false, false, false, 0, 0>);
  static OpRegistrar r13(op_key(2, 5, 1, 1, 1, 0, 1, 0, 0), &launch_op<2, 5, 1, float, true, false, true, 0, 0>);
  static FdmRegistrar f14(fdm_key(2, 6, 1, 1), &launch_fdm<2, 6, 1, float, float, float>);
  static FdmRegistrar f15(fdm_key(3, 3, 2, 1), &launch_fdm<3, 3, 2, float, float, float>);
  static OpRegistrar r16(op_key(2, 5, 1, 1, 1, 1, 0, 1, 0), &launch_op<2, 5, 1, float, true, true, false, 1, 0>);
  static OpRegistrar r17(op_key(2, 3, 3, 1, 1, 1, 0, 1, 0), &launch_op<2, 3, 3, float, true, true, false, 1, 0>);
  static OpRegistrar r18(op_key(3, 2, 4, 1, 0, 1, 0, 1, 0), &launch_op<3, 2, 4, float, false, true, false, 1, 0>);
  static OpRegistrar r19(op_key(3, 2, 3, 1, 1, 1, 0, 1, 0), &launch_op<3, 2, 3, float, true, true, false, 1, 0>);
  static OpRegistrar r20(op_key(2, 3, 1, 0, 0, 0, 0, 0, 0), &launch_op<2, 3, 1, double, false, false, false, 0, 0>);
  static OpRegistrar r21(op_key(3, 2, 1, 0, 0, 0, 0, 0, 0), &launch_op<3, 2, 1, double, false, false, false, 0, 0>);
  static FdmRegistrar f22(fdm_key(3, 2, 1, 1), &launch_fdm<3, 2, 1, float, float, float>);
} // namespace dgops
