#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 3, 1), &launch_fdm<3, 8, 3, float, float, float>);
  static OpRegistrar r1(op_key(3, 6, 4, 1, 1, 1, 1, 0), &launch_op<3, 6, 4, float, true, true, 1, 0>);
  static OpRegistrar r2(op_key(3, 8, 1, 1, 1, 1, 1, 0), &launch_op<3, 8, 1, float, true, true, 1, 0>);
  static OpRegistrar r3(op_key(3, 5, 4, 1, 1, 1, 1, 0), &launch_op<3, 5, 4, float, true, true, 1, 0>);
  static FdmRegistrar f4(fdm_key(2, 8, 4, 0), &launch_fdm<2, 8, 4, double, double, double>);
  static FdmRegistrar f5(fdm_key(2, 7, 4, 1), &launch_fdm<2, 7, 4, float, float, float>);
  static OpRegistrar r6(op_key(3, 5, 2, 1, 0, 1, 1, 0), &launch_op<3, 5, 2, float, false, true, 1, 0>);
  static FdmRegistrar f7(fdm_key(3, 4, 4, 2), &launch_fdm<3, 4, 4, float, double, double>);
  static OpRegistrar r8(op_key(2, 6, 3, 0, 1, 1, 1, 0), &launch_op<2, 6, 3, double, true, true, 1, 0>);
  static OpRegistrar r9(op_key(2, 8, 1, 0, 0, 0, 0, 0), &launch_op<2, 8, 1, double, false, false, 0, 0>);
  static OpRegistrar r10(op_key(2, 5, 4, 0, 1, 1, 1, 0), &launch_op<2, 5, 4, double, true, true, 1, 0>);
  static OpRegistrar r11(op_key(2, 7, 1, 0, 0, 0, 0, 1), &launch_op<2, 7, 1, double, false, false, 0, 1>);
  static OpRegistrar r12(op_key(3, 4, 1, 0, 0, 0, 0, 1), &launch_op<3, 4, 1, double, false, false, 0, 1>);
  static OpRegistrar r13(op_key(2, 5, 2, 1, 1, 1, 1, 0), &launch_op<2, 5, 2, float, true, true, 1, 0>);
  static OpRegistrar r14(op_key(2, 6, 1, 1, 1, 1, 1, 0), &launch_op<2, 6, 1, float, true, true, 1, 0>);
  static FdmRegistrar f15(fdm_key(2, 4, 2, 0), &launch_fdm<2, 4, 2, double, double, double>);
  static FdmRegistrar f16(fdm_key(2, 3, 4, 0), &launch_fdm<2, 3, 4, double, double, double>);
  static FdmRegistrar f17(fdm_key(3, 3, 1, 0), &launch_fdm<3, 3, 1, double, double, double>);
  static FdmRegistrar f18(fdm_key(3, 2, 4, 0), &launch_fdm<3, 2, 4, double, double, double>);
  static OpRegistrar r19(op_key(3, 2, 2, 1, 1, 1, 1, 0), &launch_op<3, 2, 2, float, true, true, 1, 0>);
  static FdmRegistrar f20(fdm_key(2, 3, 1, 1), &launch_fdm<2, 3, 1, float, float, float>);
  static OpRegistrar r21(op_key(2, 2, 2, 1, 1, 1, 1, 0), &launch_op<2, 2, 2, float, true, true, 1, 0>);
  static FdmRegistrar f22(fdm_key(2, 2, 1, 1), &launch_fdm<2, 2, 1, float, float, float>);
} // namespace dgops
