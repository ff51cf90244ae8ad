#include "../registry.hpp"

namespace dgops
{
  static BurgersRegistrar g0(burgers_key(3, 5, 0, 0), &launch_burgers<3, 5, 10, false, 0>);
  static OpRegistrar r1(op_key(3, 6, 4, 1, 1, 1, 0, 1, 0), &launch_op<3, 6, 4, float, true, true, false, 1, 0>);
  static FdmRegistrar f2(fdm_key(3, 8, 1, 0), &launch_fdm<3, 8, 1, double, double, double>);
  static FdmRegistrar f3(fdm_key(3, 5, 4, 1), &launch_fdm<3, 5, 4, float, float, float>);
  static FdmRegistrar f4(fdm_key(2, 8, 3, 0), &launch_fdm<2, 8, 3, double, double, double>);
  static OpRegistrar r5(op_key(3, 5, 2, 0, 1, 1, 0, 1, 0), &launch_op<3, 5, 2, double, true, true, false, 1, 0>);
  static OpRegistrar r6(op_key(2, 8, 1, 1, 1, 0, 1, 0, 1), &launch_op<2, 8, 1, float, true, false, true, 0, 1>);
  static OpRegistrar r7(op_key(2, 7, 1, 0, 1, 0, 1, 0, 0), &launch_op<2, 7, 1, double, true, false, true, 0, 0>);
  static FdmRegistrar f8(fdm_key(3, 5, 1, 1), &launch_fdm<3, 5, 1, float, float, float>);
  static FdmRegistrar f9(fdm_key(3, 4, 2, 2), &launch_fdm<3, 4, 2, float, double, double>);
  static OpRegistrar r10(op_key(2, 7, 1, 0, 1, 0, 0, 0, 0), &launch_op<2, 7, 1, double, true, false, false, 0, 0>);
  static OpRegistrar r11(op_key(2, 4, 4, 0, 1, 1, 0, 1, 0), &launch_op<2, 4, 4, double, true, true, false, 1, 0>);
  static OpRegistrar r12(op_key(2, 6, 1, 0, 0, 0, 0, 0, 0), &launch_op<2, 6, 1, double, false, false, false, 0, 0>);
  static OpRegistrar r13(op_key(3, 3, 1, 0, 1, 0, 1, 0, 1), &launch_op<3, 3, 1, double, true, false, true, 0, 1>);
  static OpRegistrar r14(op_key(2, 5, 1, 0, 0, 1, 0, 1, 0), &launch_op<2, 5, 1, double, false, true, false, 1, 0>);
  static FdmRegistrar f15(fdm_key(2, 3, 3, 1), &launch_fdm<2, 3, 3, float, float, float>);
  static FdmRegistrar f16(fdm_key(3, 2, 4, 1), &launch_fdm<3, 2, 4, float, float, float>);
  static OpRegistrar r17(op_key(3, 2, 1, 0, 1, 1, 1, 1, 0), &launch_op<3, 2, 1, double, true, true, true, 1, 0>);
  static BurgersRegistrar g18(burgers_key(2, 3, 1, 0), &launch_burgers<2, 3, 3, true, 0>);
  static FdmRegistrar f19(fdm_key(2, 2, 2, 0), &launch_fdm<2, 2, 2, double, double, double>);
  static BurgersRegistrar g20(burgers_key(3, 2, 1, 0), &launch_burgers<3, 2, 1, true, 0>);
} // namespace dgops
