#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 7, 4, 0, 1, 1, 1, 0), &launch_op<3, 7, 4, double, true, true, 1, 0>);
  static FdmRegistrar f1(fdm_key(3, 6, 4, 1), &launch_fdm<3, 6, 4, float, float, float>);
  static FdmRegistrar f2(fdm_key(3, 8, 1, 1), &launch_fdm<3, 8, 1, float, float, float>);
  static FdmRegistrar f3(fdm_key(3, 5, 4, 0), &launch_fdm<3, 5, 4, double, double, double>);
  static FdmRegistrar f4(fdm_key(2, 8, 4, 2), &launch_fdm<2, 8, 4, float, double, double>);
  static OpRegistrar r5(op_key(3, 6, 1, 0, 0, 0, 0, 0), &launch_op<3, 6, 1, double, false, false, 0, 0>);
  static FdmRegistrar f6(fdm_key(3, 5, 2, 1), &launch_fdm<3, 5, 2, float, float, float>);
  static OpRegistrar r7(op_key(2, 6, 4, 1, 1, 1, 1, 0), &launch_op<2, 6, 4, float, true, true, 1, 0>);
  static FdmRegistrar f8(fdm_key(2, 6, 3, 0), &launch_fdm<2, 6, 3, double, double, double>);
  static OpRegistrar r9(op_key(2, 8, 1, 0, 1, 0, 0, 0), &launch_op<2, 8, 1, double, true, false, 0, 0>);
  static FdmRegistrar f10(fdm_key(2, 5, 4, 0), &launch_fdm<2, 5, 4, double, double, double>);
  static OpRegistrar r11(op_key(2, 7, 1, 0, 1, 0, 0, 1), &launch_op<2, 7, 1, double, true, false, 0, 1>);
  static OpRegistrar r12(op_key(3, 4, 1, 0, 1, 0, 0, 1), &launch_op<3, 4, 1, double, true, false, 0, 1>);
  static FdmRegistrar f13(fdm_key(2, 5, 2, 1), &launch_fdm<2, 5, 2, float, float, float>);
  static FdmRegistrar f14(fdm_key(2, 6, 1, 1), &launch_fdm<2, 6, 1, float, float, float>);
  static FdmRegistrar f15(fdm_key(2, 4, 2, 2), &launch_fdm<2, 4, 2, float, double, double>);
  static FdmRegistrar f16(fdm_key(2, 3, 4, 2), &launch_fdm<2, 3, 4, float, double, double>);
  static FdmRegistrar f17(fdm_key(3, 3, 1, 2), &launch_fdm<3, 3, 1, float, double, double>);
  static FdmRegistrar f18(fdm_key(3, 2, 4, 2), &launch_fdm<3, 2, 4, float, double, double>);
  static FdmRegistrar f19(fdm_key(2, 2, 4, 1), &launch_fdm<2, 2, 4, float, float, float>);
  static OpRegistrar r20(op_key(2, 2, 3, 0, 1, 1, 1, 0), &launch_op<2, 2, 3, double, true, true, 1, 0>);
  static FdmRegistrar f21(fdm_key(2, 2, 2, 1), &launch_fdm<2, 2, 2, float, float, float>);
} // namespace dgops
