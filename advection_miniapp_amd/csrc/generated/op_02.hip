#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 4, 2), &launch_fdm<3, 8, 4, float, double, double>);
  static FdmRegistrar f1(fdm_key(3, 7, 3, 2), &launch_fdm<3, 7, 3, float, double, double>);
  static OpRegistrar r2(op_key(3, 8, 1, 1, 0, 1, 1, 0), &launch_op<3, 8, 1, float, false, true, 1, 0>);
  static OpRegistrar r3(op_key(3, 5, 4, 0, 1, 1, 1, 0), &launch_op<3, 5, 4, double, true, true, 1, 0>);
  static FdmRegistrar f4(fdm_key(3, 7, 1, 1), &launch_fdm<3, 7, 1, float, float, float>);
  static OpRegistrar r5(op_key(2, 7, 4, 1, 1, 1, 1, 0), &launch_op<2, 7, 4, float, true, true, 1, 0>);
  static OpRegistrar r6(op_key(3, 5, 2, 0, 0, 1, 1, 0), &launch_op<3, 5, 2, double, false, true, 1, 0>);
  static FdmRegistrar f7(fdm_key(3, 4, 4, 0), &launch_fdm<3, 4, 4, double, double, double>);
  static FdmRegistrar f8(fdm_key(2, 7, 2, 1), &launch_fdm<2, 7, 2, float, float, float>);
  static FdmRegistrar f9(fdm_key(3, 5, 1, 1), &launch_fdm<3, 5, 1, float, float, float>);
  static FdmRegistrar f10(fdm_key(3, 4, 2, 1), &launch_fdm<3, 4, 2, float, float, float>);
  static FdmRegistrar f11(fdm_key(2, 5, 3, 2), &launch_fdm<2, 5, 3, float, double, double>);
  static FdmRegistrar f12(fdm_key(3, 3, 4, 2), &launch_fdm<3, 3, 4, float, double, double>);
  static FdmRegistrar f13(fdm_key(3, 4, 1, 2), &launch_fdm<3, 4, 1, float, double, double>);
  static OpRegistrar r14(op_key(2, 6, 1, 1, 0, 1, 1, 0), &launch_op<2, 6, 1, float, false, true, 1, 0>);
  static OpRegistrar r15(op_key(2, 4, 2, 0, 1, 1, 1, 0), &launch_op<2, 4, 2, double, true, true, 1, 0>);
  static OpRegistrar r16(op_key(2, 3, 4, 0, 1, 1, 1, 0), &launch_op<2, 3, 4, double, true, true, 1, 0>);
  static FdmRegistrar f17(fdm_key(2, 3, 3, 1), &launch_fdm<2, 3, 3, float, float, float>);
  static FdmRegistrar f18(fdm_key(2, 4, 1, 1), &launch_fdm<2, 4, 1, float, float, float>);
  static OpRegistrar r19(op_key(2, 2, 4, 1, 1, 1, 1, 0), &launch_op<2, 2, 4, float, true, true, 1, 0>);
  static OpRegistrar r20(op_key(2, 3, 1, 1, 1, 1, 1, 0), &launch_op<2, 3, 1, float, true, true, 1, 0>);
  static OpRegistrar r21(op_key(3, 2, 1, 1, 1, 1, 1, 0), &launch_op<3, 2, 1, float, true, true, 1, 0>);
  static OpRegistrar r22(op_key(2, 2, 1, 1, 1, 1, 1, 0), &launch_op<2, 2, 1, float, true, true, 1, 0>);
} // namespace dgops
