#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 4, 1), &launch_fdm<3, 8, 4, float, float, float>);
  static FdmRegistrar f1(fdm_key(3, 7, 3, 1), &launch_fdm<3, 7, 3, float, float, float>);
  static OpRegistrar r2(op_key(3, 8, 1, 0, 0, 1, 1, 0), &launch_op<3, 8, 1, double, false, true, 1, 0>);
  static FdmRegistrar f3(fdm_key(3, 6, 2, 2), &launch_fdm<3, 6, 2, float, double, double>);
  static FdmRegistrar f4(fdm_key(3, 7, 1, 0), &launch_fdm<3, 7, 1, double, double, double>);
  static OpRegistrar r5(op_key(2, 7, 4, 0, 1, 1, 1, 0), &launch_op<2, 7, 4, double, true, true, 1, 0>);
  static OpRegistrar r6(op_key(3, 5, 2, 0, 1, 1, 1, 0), &launch_op<3, 5, 2, double, true, true, 1, 0>);
  static FdmRegistrar f7(fdm_key(2, 8, 2, 2), &launch_fdm<2, 8, 2, float, double, double>);
  static FdmRegistrar f8(fdm_key(2, 7, 2, 0), &launch_fdm<2, 7, 2, double, double, double>);
  static FdmRegistrar f9(fdm_key(3, 5, 1, 0), &launch_fdm<3, 5, 1, double, double, double>);
  static FdmRegistrar f10(fdm_key(3, 4, 2, 0), &launch_fdm<3, 4, 2, double, double, double>);
  static FdmRegistrar f11(fdm_key(2, 5, 3, 1), &launch_fdm<2, 5, 3, float, float, float>);
  static FdmRegistrar f12(fdm_key(3, 3, 4, 1), &launch_fdm<3, 3, 4, float, float, float>);
  static FdmRegistrar f13(fdm_key(3, 4, 1, 1), &launch_fdm<3, 4, 1, float, float, float>);
  static OpRegistrar r14(op_key(2, 6, 1, 0, 0, 1, 1, 0), &launch_op<2, 6, 1, double, false, true, 1, 0>);
  static FdmRegistrar f15(fdm_key(3, 3, 2, 2), &launch_fdm<3, 3, 2, float, double, double>);
  static FdmRegistrar f16(fdm_key(2, 5, 1, 2), &launch_fdm<2, 5, 1, float, double, double>);
  static FdmRegistrar f17(fdm_key(2, 3, 3, 0), &launch_fdm<2, 3, 3, double, double, double>);
  static FdmRegistrar f18(fdm_key(2, 4, 1, 0), &launch_fdm<2, 4, 1, double, double, double>);
  static OpRegistrar r19(op_key(2, 2, 4, 0, 1, 1, 1, 0), &launch_op<2, 2, 4, double, true, true, 1, 0>);
  static OpRegistrar r20(op_key(2, 3, 1, 0, 1, 1, 1, 0), &launch_op<2, 3, 1, double, true, true, 1, 0>);
  static OpRegistrar r21(op_key(3, 2, 1, 0, 1, 1, 1, 0), &launch_op<3, 2, 1, double, true, true, 1, 0>);
  static OpRegistrar r22(op_key(2, 2, 1, 0, 1, 1, 1, 0), &launch_op<2, 2, 1, double, true, true, 1, 0>);
} // namespace dgops
