#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 4, 0), &launch_fdm<3, 8, 4, double, double, double>);
  static FdmRegistrar f1(fdm_key(3, 7, 3, 0), &launch_fdm<3, 7, 3, double, double, double>);
  static OpRegistrar r2(op_key(3, 8, 1, 0, 1, 0, 0, 1), &launch_op<3, 8, 1, double, true, false, 0, 1>);
  static FdmRegistrar f3(fdm_key(3, 6, 2, 1), &launch_fdm<3, 6, 2, float, float, float>);
  static OpRegistrar r4(op_key(3, 7, 1, 1, 1, 1, 1, 0), &launch_op<3, 7, 1, float, true, true, 1, 0>);
  static FdmRegistrar f5(fdm_key(2, 8, 3, 2), &launch_fdm<2, 8, 3, float, double, double>);
  static FdmRegistrar f6(fdm_key(3, 6, 1, 2), &launch_fdm<3, 6, 1, float, double, double>);
  static FdmRegistrar f7(fdm_key(2, 8, 2, 1), &launch_fdm<2, 8, 2, float, float, float>);
  static OpRegistrar r8(op_key(2, 7, 2, 1, 1, 1, 1, 0), &launch_op<2, 7, 2, float, true, true, 1, 0>);
  static OpRegistrar r9(op_key(3, 5, 1, 1, 1, 1, 1, 0), &launch_op<3, 5, 1, float, true, true, 1, 0>);
  static FdmRegistrar f10(fdm_key(2, 8, 1, 2), &launch_fdm<2, 8, 1, float, double, double>);
  static FdmRegistrar f11(fdm_key(2, 5, 3, 0), &launch_fdm<2, 5, 3, double, double, double>);
  static FdmRegistrar f12(fdm_key(3, 3, 4, 0), &launch_fdm<3, 3, 4, double, double, double>);
  static FdmRegistrar f13(fdm_key(3, 4, 1, 0), &launch_fdm<3, 4, 1, double, double, double>);
  static OpRegistrar r14(op_key(2, 6, 1, 0, 1, 0, 0, 1), &launch_op<2, 6, 1, double, true, false, 0, 1>);
  static FdmRegistrar f15(fdm_key(3, 3, 2, 1), &launch_fdm<3, 3, 2, float, float, float>);
  static FdmRegistrar f16(fdm_key(2, 5, 1, 1), &launch_fdm<2, 5, 1, float, float, float>);
  static OpRegistrar r17(op_key(2, 3, 3, 1, 1, 1, 1, 0), &launch_op<2, 3, 3, float, true, true, 1, 0>);
  static OpRegistrar r18(op_key(3, 2, 4, 1, 1, 1, 1, 0), &launch_op<3, 2, 4, float, true, true, 1, 0>);
  static FdmRegistrar f19(fdm_key(3, 2, 3, 2), &launch_fdm<3, 2, 3, float, double, double>);
  static OpRegistrar r20(op_key(2, 3, 1, 1, 0, 1, 1, 0), &launch_op<2, 3, 1, float, false, true, 1, 0>);
  static OpRegistrar r21(op_key(3, 2, 1, 1, 0, 1, 1, 0), &launch_op<3, 2, 1, float, false, true, 1, 0>);
  static OpRegistrar r22(op_key(2, 2, 1, 1, 0, 1, 1, 0), &launch_op<2, 2, 1, float, false, true, 1, 0>);
} // namespace dgops
