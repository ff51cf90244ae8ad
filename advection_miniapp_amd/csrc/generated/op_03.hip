#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 3, 0), &launch_fdm<3, 8, 3, double, double, double>);
  static FdmRegistrar f1(fdm_key(3, 7, 3, 1), &launch_fdm<3, 7, 3, float, float, float>);
  static OpRegistrar r2(op_key(3, 8, 1, 1, 0, 1, 0, 1, 0), &launch_op<3, 8, 1, float, false, true, false, 1, 0>);
  static OpRegistrar r3(op_key(3, 5, 4, 1, 1, 1, 0, 1, 0), &launch_op<3, 5, 4, float, true, true, false, 1, 0>);
  static FdmRegistrar f4(fdm_key(3, 5, 3, 0), &launch_fdm<3, 5, 3, double, double, double>);
  static BurgersRegistrar g5(burgers_key(3, 4, 0, 1), &launch_burgers<3, 4, 6, false, 1>);
  static OpRegistrar r6(op_key(2, 8, 1, 0, 1, 0, 1, 0, 1), &launch_op<2, 8, 1, double, true, false, true, 0, 1>);
  static FdmRegistrar f7(fdm_key(3, 4, 3, 2), &launch_fdm<3, 4, 3, float, double, double>);
  static OpRegistrar r8(op_key(3, 5, 1, 0, 1, 1, 0, 1, 0), &launch_op<3, 5, 1, double, true, true, false, 1, 0>);
  static FdmRegistrar f9(fdm_key(2, 8, 1, 2), &launch_fdm<2, 8, 1, float, double, double>);
  static FdmRegistrar f10(fdm_key(2, 5, 3, 2), &launch_fdm<2, 5, 3, float, double, double>);
  static OpRegistrar r11(op_key(3, 4, 1, 1, 0, 1, 0, 1, 0), &launch_op<3, 4, 1, float, false, true, false, 1, 0>);
  static FdmRegistrar f12(fdm_key(3, 3, 3, 0), &launch_fdm<3, 3, 3, double, double, double>);
  static OpRegistrar r13(op_key(3, 3, 2, 0, 1, 1, 0, 1, 0), &launch_op<3, 3, 2, double, true, true, false, 1, 0>);
  static OpRegistrar r14(op_key(2, 5, 1, 0, 1, 0, 0, 0, 0), &launch_op<2, 5, 1, double, true, false, false, 0, 0>);
  static OpRegistrar r15(op_key(2, 3, 3, 0, 1, 1, 0, 1, 0), &launch_op<2, 3, 3, double, true, true, false, 1, 0>);
  static FdmRegistrar f16(fdm_key(2, 4, 1, 1), &launch_fdm<2, 4, 1, float, float, float>);
  static OpRegistrar r17(op_key(3, 2, 2, 1, 1, 1, 0, 1, 0), &launch_op<3, 2, 2, float, true, true, false, 1, 0>);
  static FdmRegistrar f18(fdm_key(2, 3, 1, 2), &launch_fdm<2, 3, 1, float, double, double>);
  static OpRegistrar r19(op_key(2, 2, 1, 1, 1, 0, 1, 0, 0), &launch_op<2, 2, 1, float, true, false, true, 0, 0>);
  static BurgersRegistrar g20(burgers_key(2, 2, 1, 1), &launch_burgers<2, 2, 1, true, 1>);
} // namespace dgops
