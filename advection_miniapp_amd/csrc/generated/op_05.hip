#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 3, 2), &launch_fdm<3, 8, 3, float, double, double>);
  static OpRegistrar r1(op_key(3, 6, 4, 0, 1, 1, 0, 1, 0), &launch_op<3, 6, 4, double, true, true, false, 1, 0>);
  static OpRegistrar r2(op_key(3, 8, 1, 1, 1, 1, 0, 1, 0), &launch_op<3, 8, 1, float, true, true, false, 1, 0>);
  static FdmRegistrar f3(fdm_key(3, 5, 4, 0), &launch_fdm<3, 5, 4, double, double, double>);
  static FdmRegistrar f4(fdm_key(3, 5, 3, 2), &launch_fdm<3, 5, 3, float, double, double>);
  static BurgersRegistrar g5(burgers_key(3, 4, 1, 1), &launch_burgers<3, 4, 6, true, 1>);
  static OpRegistrar r6(op_key(2, 8, 1, 1, 1, 0, 1, 0, 0), &launch_op<2, 8, 1, float, true, false, true, 0, 0>);
  static OpRegistrar r7(op_key(2, 7, 2, 1, 1, 1, 0, 1, 0), &launch_op<2, 7, 2, float, true, true, false, 1, 0>);
  static FdmRegistrar f8(fdm_key(3, 5, 1, 0), &launch_fdm<3, 5, 1, double, double, double>);
  static FdmRegistrar f9(fdm_key(3, 4, 2, 1), &launch_fdm<3, 4, 2, float, float, float>);
  static OpRegistrar r10(op_key(2, 7, 1, 0, 0, 0, 0, 0, 1), &launch_op<2, 7, 1, double, false, false, false, 0, 1>);
  static OpRegistrar r11(op_key(3, 4, 1, 1, 1, 1, 0, 1, 0), &launch_op<3, 4, 1, float, true, true, false, 1, 0>);
  static FdmRegistrar f12(fdm_key(3, 3, 3, 2), &launch_fdm<3, 3, 3, float, double, double>);
  static OpRegistrar r13(op_key(3, 3, 1, 0, 1, 0, 1, 0, 0), &launch_op<3, 3, 1, double, true, false, true, 0, 0>);
  static OpRegistrar r14(op_key(2, 5, 1, 1, 1, 0, 0, 0, 0), &launch_op<2, 5, 1, float, true, false, false, 0, 0>);
  static FdmRegistrar f15(fdm_key(2, 3, 3, 0), &launch_fdm<2, 3, 3, double, double, double>);
  static FdmRegistrar f16(fdm_key(3, 2, 4, 0), &launch_fdm<3, 2, 4, double, double, double>);
  static OpRegistrar r17(op_key(3, 2, 1, 0, 1, 0, 1, 0, 1), &launch_op<3, 2, 1, double, true, false, true, 0, 1>);
  static BurgersRegistrar g18(burgers_key(2, 3, 0, 1), &launch_burgers<2, 3, 3, false, 1>);
  static OpRegistrar r19(op_key(2, 2, 1, 1, 1, 1, 1, 1, 0), &launch_op<2, 2, 1, float, true, true, true, 1, 0>);
  static BurgersRegistrar g20(burgers_key(3, 2, 0, 1), &launch_burgers<3, 2, 1, false, 1>);
} // namespace dgops
