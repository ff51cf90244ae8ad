#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 8, 1, 0, 1, 0, 1, 0, 0), &launch_op<3, 8, 1, double, true, false, true, 0, 0>);
  static OpRegistrar r1(op_key(3, 7, 2, 0, 1, 1, 0, 1, 0), &launch_op<3, 7, 2, double, true, true, false, 1, 0>);
  static OpRegistrar r2(op_key(3, 6, 3, 0, 0, 1, 0, 1, 0), &launch_op<3, 6, 3, double, false, true, false, 1, 0>);
  static FdmRegistrar f3(fdm_key(3, 5, 4, 0), &launch_fdm<3, 5, 4, double, double, double>);
  static FdmRegistrar f4(fdm_key(3, 5, 3, 2), &launch_fdm<3, 5, 3, float, double, double>);
  static BurgersRegistrar g5(burgers_key(3, 4, 0, 1), &launch_burgers<3, 4, 6, false, 1>);
  static OpRegistrar r6(op_key(3, 4, 4, 0, 1, 1, 0, 1, 0), &launch_op<3, 4, 4, double, true, true, false, 1, 0>);
  static FdmRegistrar f7(fdm_key(2, 6, 4, 1), &launch_fdm<2, 6, 4, float, float, float>);
  static OpRegistrar r8(op_key(2, 6, 3, 1, 0, 1, 0, 1, 0), &launch_op<2, 6, 3, float, false, true, false, 1, 0>);
  static OpRegistrar r9(op_key(2, 8, 1, 0, 1, 1, 0, 1, 0), &launch_op<2, 8, 1, double, true, true, false, 1, 0>);
  static FdmRegistrar f10(fdm_key(2, 5, 4, 2), &launch_fdm<2, 5, 4, float, double, double>);
  static OpRegistrar r11(op_key(2, 7, 1, 0, 1, 0, 0, 0, 1), &launch_op<2, 7, 1, double, true, false, false, 0, 1>);
  static OpRegistrar r12(op_key(3, 4, 1, 1, 1, 1, 0, 1, 0), &launch_op<3, 4, 1, float, true, true, false, 1, 0>);
  static OpRegistrar r13(op_key(3, 3, 3, 0, 0, 1, 0, 1, 0), &launch_op<3, 3, 3, double, false, true, false, 1, 0>);
  static OpRegistrar r14(op_key(2, 4, 3, 0, 1, 1, 0, 1, 0), &launch_op<2, 4, 3, double, true, true, false, 1, 0>);
  static OpRegistrar r15(op_key(2, 4, 1, 0, 1, 0, 1, 0, 1), &launch_op<2, 4, 1, double, true, false, true, 0, 1>);
  static FdmRegistrar f16(fdm_key(2, 3, 4, 0), &launch_fdm<2, 3, 4, double, double, double>);
  static BurgersRegistrar g17(burgers_key(3, 3, 1, 0), &launch_burgers<3, 3, 3, true, 0>);
  static OpRegistrar r18(op_key(2, 3, 2, 1, 1, 1, 0, 1, 0), &launch_op<2, 3, 2, float, true, true, false, 1, 0>);
  static OpRegistrar r19(op_key(3, 2, 2, 1, 1, 1, 0, 1, 0), &launch_op<3, 2, 2, float, true, true, false, 1, 0>);
  static FdmRegistrar f20(fdm_key(2, 3, 1, 2), &launch_fdm<2, 3, 1, float, double, double>);
  static OpRegistrar r21(op_key(2, 2, 2, 1, 1, 1, 0, 1, 0), &launch_op<2, 2, 2, float, true, true, false, 1, 0>);
  static FdmRegistrar f22(fdm_key(2, 2, 1, 2), &launch_fdm<2, 2, 1, float, double, double>);
} // namespace dgops
