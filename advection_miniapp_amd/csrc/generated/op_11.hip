#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 2, 0), &launch_fdm<3, 8, 2, double, double, double>);
  static FdmRegistrar f1(fdm_key(3, 7, 2, 1), &launch_fdm<3, 7, 2, float, float, float>);
  static FdmRegistrar f2(fdm_key(3, 6, 3, 1), &launch_fdm<3, 6, 3, float, float, float>);
  static OpRegistrar r3(op_key(3, 7, 1, 0, 1, 0, 0, 0), &launch_op<3, 7, 1, double, true, false, 0, 0>);
  static FdmRegistrar f4(fdm_key(3, 5, 3, 0), &launch_fdm<3, 5, 3, double, double, double>);
  static OpRegistrar r5(op_key(3, 6, 1, 1, 0, 1, 1, 0), &launch_op<3, 6, 1, float, false, true, 1, 0>);
  static FdmRegistrar f6(fdm_key(2, 7, 3, 1), &launch_fdm<2, 7, 3, float, float, float>);
  static OpRegistrar r7(op_key(3, 4, 3, 1, 1, 1, 1, 0), &launch_op<3, 4, 3, float, true, true, 1, 0>);
  static OpRegistrar r8(op_key(3, 5, 1, 0, 1, 0, 0, 0), &launch_op<3, 5, 1, double, true, false, 0, 0>);
  static OpRegistrar r9(op_key(2, 8, 1, 1, 1, 1, 1, 0), &launch_op<2, 8, 1, float, true, true, 1, 0>);
  static FdmRegistrar f10(fdm_key(2, 6, 2, 0), &launch_fdm<2, 6, 2, double, double, double>);
  static FdmRegistrar f11(fdm_key(2, 7, 1, 0), &launch_fdm<2, 7, 1, double, double, double>);
  static OpRegistrar r12(op_key(2, 4, 4, 0, 1, 1, 1, 0), &launch_op<2, 4, 4, double, true, true, 1, 0>);
  static FdmRegistrar f13(fdm_key(3, 3, 3, 1), &launch_fdm<3, 3, 3, float, float, float>);
  static FdmRegistrar f14(fdm_key(2, 4, 3, 1), &launch_fdm<2, 4, 3, float, float, float>);
  static OpRegistrar r15(op_key(2, 5, 1, 0, 0, 1, 1, 0), &launch_op<2, 5, 1, double, false, true, 1, 0>);
  static OpRegistrar r16(op_key(3, 3, 1, 0, 0, 1, 1, 0), &launch_op<3, 3, 1, double, false, true, 1, 0>);
  static OpRegistrar r17(op_key(2, 4, 1, 0, 0, 1, 1, 0), &launch_op<2, 4, 1, double, false, true, 1, 0>);
  static FdmRegistrar f18(fdm_key(2, 3, 2, 2), &launch_fdm<2, 3, 2, float, double, double>);
  static OpRegistrar r19(op_key(2, 3, 1, 0, 0, 0, 0, 0), &launch_op<2, 3, 1, double, false, false, 0, 0>);
  static OpRegistrar r20(op_key(3, 2, 1, 0, 0, 0, 0, 0), &launch_op<3, 2, 1, double, false, false, 0, 0>);
  static OpRegistrar r21(op_key(2, 2, 1, 0, 0, 0, 0, 0), &launch_op<2, 2, 1, double, false, false, 0, 0>);
} // namespace dgops
