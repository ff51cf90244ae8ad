#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 7, 4, 1, 1, 1, 1, 0), &launch_op<3, 7, 4, float, true, true, 1, 0>);
  static FdmRegistrar f1(fdm_key(3, 6, 4, 2), &launch_fdm<3, 6, 4, float, double, double>);
  static FdmRegistrar f2(fdm_key(3, 8, 1, 2), &launch_fdm<3, 8, 1, float, double, double>);
  static FdmRegistrar f3(fdm_key(3, 5, 4, 1), &launch_fdm<3, 5, 4, float, float, float>);
  static OpRegistrar r4(op_key(3, 5, 3, 0, 1, 1, 1, 0), &launch_op<3, 5, 3, double, true, true, 1, 0>);
  static OpRegistrar r5(op_key(3, 6, 1, 0, 0, 0, 0, 1), &launch_op<3, 6, 1, double, false, false, 0, 1>);
  static FdmRegistrar f6(fdm_key(3, 5, 2, 2), &launch_fdm<3, 5, 2, float, double, double>);
  static FdmRegistrar f7(fdm_key(2, 6, 4, 0), &launch_fdm<2, 6, 4, double, double, double>);
  static FdmRegistrar f8(fdm_key(2, 6, 3, 1), &launch_fdm<2, 6, 3, float, float, float>);
  static OpRegistrar r9(op_key(2, 8, 1, 0, 1, 0, 0, 1), &launch_op<2, 8, 1, double, true, false, 0, 1>);
  static FdmRegistrar f10(fdm_key(2, 5, 4, 1), &launch_fdm<2, 5, 4, float, float, float>);
  static OpRegistrar r11(op_key(2, 7, 1, 0, 0, 1, 1, 0), &launch_op<2, 7, 1, double, false, true, 1, 0>);
  static OpRegistrar r12(op_key(3, 4, 1, 0, 0, 1, 1, 0), &launch_op<3, 4, 1, double, false, true, 1, 0>);
  static FdmRegistrar f13(fdm_key(2, 5, 2, 2), &launch_fdm<2, 5, 2, float, double, double>);
  static FdmRegistrar f14(fdm_key(2, 6, 1, 2), &launch_fdm<2, 6, 1, float, double, double>);
  static OpRegistrar r15(op_key(2, 5, 1, 0, 0, 0, 0, 0), &launch_op<2, 5, 1, double, false, false, 0, 0>);
  static OpRegistrar r16(op_key(3, 3, 1, 0, 0, 0, 0, 0), &launch_op<3, 3, 1, double, false, false, 0, 0>);
  static OpRegistrar r17(op_key(2, 4, 1, 0, 0, 0, 0, 0), &launch_op<2, 4, 1, double, false, false, 0, 0>);
  static OpRegistrar r18(op_key(2, 3, 2, 0, 1, 1, 1, 0), &launch_op<2, 3, 2, double, true, true, 1, 0>);
  static FdmRegistrar f19(fdm_key(2, 2, 4, 2), &launch_fdm<2, 2, 4, float, double, double>);
  static OpRegistrar r20(op_key(2, 2, 3, 1, 1, 1, 1, 0), &launch_op<2, 2, 3, float, true, true, 1, 0>);
  static FdmRegistrar f21(fdm_key(2, 2, 2, 2), &launch_fdm<2, 2, 2, float, double, double>);
} // namespace dgops
