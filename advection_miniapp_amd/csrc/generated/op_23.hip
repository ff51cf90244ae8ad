#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 2, 0), &launch_fdm<3, 8, 2, double, double, double>);
  static OpRegistrar r1(op_key(3, 7, 1, 1, 1, 0, 1, 0, 0), &launch_op<3, 7, 1, float, true, false, true, 0, 0>);
  static FdmRegistrar f2(fdm_key(3, 6, 3, 1), &launch_fdm<3, 6, 3, float, float, float>);
  static OpRegistrar r3(op_key(3, 7, 1, 0, 1, 0, 0, 0, 0), &launch_op<3, 7, 1, double, true, false, false, 0, 0>);
  static FdmRegistrar f4(fdm_key(2, 8, 3, 2), &launch_fdm<2, 8, 3, float, double, double>);
  static BurgersRegistrar g5(burgers_key(3, 4, 1, 1), &launch_burgers<3, 4, 6, true, 1>);
  static OpRegistrar r6(op_key(3, 4, 4, 1, 0, 1, 0, 1, 0), &launch_op<3, 4, 4, float, false, true, false, 1, 0>);
  static FdmRegistrar f7(fdm_key(2, 6, 4, 2), &launch_fdm<2, 6, 4, float, double, double>);
  static OpRegistrar r8(op_key(2, 6, 3, 1, 1, 1, 0, 1, 0), &launch_op<2, 6, 3, float, true, true, false, 1, 0>);
  static OpRegistrar r9(op_key(2, 8, 1, 0, 0, 1, 0, 1, 0), &launch_op<2, 8, 1, double, false, true, false, 1, 0>);
  static FdmRegistrar f10(fdm_key(2, 5, 4, 0), &launch_fdm<2, 5, 4, double, double, double>);
  static OpRegistrar r11(op_key(2, 7, 1, 0, 0, 0, 0, 0, 1), &launch_op<2, 7, 1, double, false, false, false, 0, 1>);
  static OpRegistrar r12(op_key(3, 4, 1, 1, 1, 0, 0, 0, 0), &launch_op<3, 4, 1, float, true, false, false, 0, 0>);
  static OpRegistrar r13(op_key(2, 5, 1, 1, 1, 0, 1, 0, 1), &launch_op<2, 5, 1, float, true, false, true, 0, 1>);
  static FdmRegistrar f14(fdm_key(2, 6, 1, 2), &launch_fdm<2, 6, 1, float, double, double>);
  static FdmRegistrar f15(fdm_key(3, 3, 2, 2), &launch_fdm<3, 3, 2, float, double, double>);
  static FdmRegistrar f16(fdm_key(2, 5, 1, 0), &launch_fdm<2, 5, 1, double, double, double>);
  static FdmRegistrar f17(fdm_key(2, 3, 3, 0), &launch_fdm<2, 3, 3, double, double, double>);
  static OpRegistrar r18(op_key(3, 2, 4, 1, 1, 1, 0, 1, 0), &launch_op<3, 2, 4, float, true, true, false, 1, 0>);
  static FdmRegistrar f19(fdm_key(3, 2, 3, 0), &launch_fdm<3, 2, 3, double, double, double>);
  static OpRegistrar r20(op_key(2, 3, 1, 0, 0, 0, 0, 0, 1), &launch_op<2, 3, 1, double, false, false, false, 0, 1>);
  static OpRegistrar r21(op_key(3, 2, 1, 0, 0, 0, 0, 0, 1), &launch_op<3, 2, 1, double, false, false, false, 0, 1>);
  static FdmRegistrar f22(fdm_key(3, 2, 1, 2), &launch_fdm<3, 2, 1, float, double, double>);
} // namespace dgops
