#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 7, 4, 0), &launch_fdm<3, 7, 4, double, double, double>);
  static OpRegistrar r1(op_key(3, 7, 2, 0, 1, 1, 1, 0), &launch_op<3, 7, 2, double, true, true, 1, 0>);
  static OpRegistrar r2(op_key(3, 6, 3, 0, 1, 1, 1, 0), &launch_op<3, 6, 3, double, true, true, 1, 0>);
  static FdmRegistrar f3(fdm_key(3, 5, 4, 2), &launch_fdm<3, 5, 4, float, double, double>);
  static OpRegistrar r4(op_key(3, 5, 3, 0, 0, 1, 1, 0), &launch_op<3, 5, 3, double, false, true, 1, 0>);
  static OpRegistrar r5(op_key(3, 6, 1, 0, 1, 0, 0, 0), &launch_op<3, 6, 1, double, true, false, 0, 0>);
  static OpRegistrar r6(op_key(2, 7, 3, 0, 1, 1, 1, 0), &launch_op<2, 7, 3, double, true, true, 1, 0>);
  static FdmRegistrar f7(fdm_key(2, 6, 4, 1), &launch_fdm<2, 6, 4, float, float, float>);
  static FdmRegistrar f8(fdm_key(2, 6, 3, 2), &launch_fdm<2, 6, 3, float, double, double>);
  static OpRegistrar r9(op_key(2, 8, 1, 0, 0, 1, 1, 0), &launch_op<2, 8, 1, double, false, true, 1, 0>);
  static FdmRegistrar f10(fdm_key(2, 5, 4, 2), &launch_fdm<2, 5, 4, float, double, double>);
  static OpRegistrar r11(op_key(2, 7, 1, 1, 0, 1, 1, 0), &launch_op<2, 7, 1, float, false, true, 1, 0>);
  static OpRegistrar r12(op_key(3, 4, 1, 1, 0, 1, 1, 0), &launch_op<3, 4, 1, float, false, true, 1, 0>);
  static OpRegistrar r13(op_key(3, 3, 3, 0, 1, 1, 1, 0), &launch_op<3, 3, 3, double, true, true, 1, 0>);
  static OpRegistrar r14(op_key(2, 4, 3, 0, 1, 1, 1, 0), &launch_op<2, 4, 3, double, true, true, 1, 0>);
  static OpRegistrar r15(op_key(2, 5, 1, 0, 0, 0, 0, 1), &launch_op<2, 5, 1, double, false, false, 0, 1>);
  static OpRegistrar r16(op_key(3, 3, 1, 0, 0, 0, 0, 1), &launch_op<3, 3, 1, double, false, false, 0, 1>);
  static OpRegistrar r17(op_key(2, 4, 1, 0, 0, 0, 0, 1), &launch_op<2, 4, 1, double, false, false, 0, 1>);
  static OpRegistrar r18(op_key(2, 3, 2, 1, 1, 1, 1, 0), &launch_op<2, 3, 2, float, true, true, 1, 0>);
  static FdmRegistrar f19(fdm_key(3, 2, 2, 0), &launch_fdm<3, 2, 2, double, double, double>);
  static FdmRegistrar f20(fdm_key(2, 2, 3, 0), &launch_fdm<2, 2, 3, double, double, double>);
  static FdmRegistrar f21(fdm_key(3, 2, 1, 0), &launch_fdm<3, 2, 1, double, double, double>);
} // namespace dgops
