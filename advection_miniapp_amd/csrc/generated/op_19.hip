#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 8, 1, 0, 1, 1, 1, 1, 0), &launch_op<3, 8, 1, double, true, true, true, 1, 0>);
  static OpRegistrar r1(op_key(3, 7, 2, 1, 1, 1, 0, 1, 0), &launch_op<3, 7, 2, float, true, true, false, 1, 0>);
  static OpRegistrar r2(op_key(3, 6, 3, 0, 1, 1, 0, 1, 0), &launch_op<3, 6, 3, double, true, true, false, 1, 0>);
  static FdmRegistrar f3(fdm_key(3, 5, 4, 1), &launch_fdm<3, 5, 4, float, float, float>);
  static FdmRegistrar f4(fdm_key(3, 5, 3, 1), &launch_fdm<3, 5, 3, float, float, float>);
  static FdmRegistrar f5(fdm_key(3, 6, 1, 2), &launch_fdm<3, 6, 1, float, double, double>);
  static FdmRegistrar f6(fdm_key(2, 7, 3, 1), &launch_fdm<2, 7, 3, float, float, float>);
  static OpRegistrar r7(op_key(2, 6, 4, 1, 0, 1, 0, 1, 0), &launch_op<2, 6, 4, float, false, true, false, 1, 0>);
  static FdmRegistrar f8(fdm_key(2, 7, 2, 2), &launch_fdm<2, 7, 2, float, double, double>);
  static OpRegistrar r9(op_key(2, 8, 1, 0, 1, 0, 0, 0, 0), &launch_op<2, 8, 1, double, true, false, false, 0, 0>);
  static OpRegistrar r10(op_key(2, 5, 4, 0, 0, 1, 0, 1, 0), &launch_op<2, 5, 4, double, false, true, false, 1, 0>);
  static FdmRegistrar f11(fdm_key(2, 5, 3, 0), &launch_fdm<2, 5, 3, double, double, double>);
  static OpRegistrar r12(op_key(3, 4, 1, 0, 0, 0, 0, 0, 1), &launch_op<3, 4, 1, double, false, false, false, 0, 1>);
  static OpRegistrar r13(op_key(2, 5, 1, 0, 1, 0, 1, 0, 0), &launch_op<2, 5, 1, double, true, false, true, 0, 0>);
  static OpRegistrar r14(op_key(2, 6, 1, 0, 1, 1, 0, 1, 0), &launch_op<2, 6, 1, double, true, true, false, 1, 0>);
  static OpRegistrar r15(op_key(3, 3, 1, 1, 1, 0, 1, 0, 1), &launch_op<3, 3, 1, float, true, false, true, 0, 1>);
  static OpRegistrar r16(op_key(2, 5, 1, 0, 0, 1, 0, 1, 0), &launch_op<2, 5, 1, double, false, true, false, 1, 0>);
  static OpRegistrar r17(op_key(2, 3, 3, 0, 0, 1, 0, 1, 0), &launch_op<2, 3, 3, double, false, true, false, 1, 0>);
  static OpRegistrar r18(op_key(2, 4, 1, 1, 1, 1, 0, 1, 0), &launch_op<2, 4, 1, float, true, true, false, 1, 0>);
  static OpRegistrar r19(op_key(3, 2, 3, 0, 0, 1, 0, 1, 0), &launch_op<3, 2, 3, double, false, true, false, 1, 0>);
  static FdmRegistrar f20(fdm_key(3, 2, 2, 0), &launch_fdm<3, 2, 2, double, double, double>);
  static FdmRegistrar f21(fdm_key(2, 2, 3, 0), &launch_fdm<2, 2, 3, double, double, double>);
  static FdmRegistrar f22(fdm_key(2, 2, 2, 1), &launch_fdm<2, 2, 2, float, float, float>);
  static BurgersRegistrar g23(burgers_key(3, 2, 1, 0), &launch_burgers<3, 2, 1, true, 0>);
} // namespace dgops
