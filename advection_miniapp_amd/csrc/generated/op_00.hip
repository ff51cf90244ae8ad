#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 4, 0), &launch_fdm<3, 8, 4, double, double, double>);
  static FdmRegistrar f1(fdm_key(3, 8, 2, 1), &launch_fdm<3, 8, 2, float, float, float>);
  static OpRegistrar r2(op_key(3, 7, 1, 1, 1, 0, 1, 0, 1), &launch_op<3, 7, 1, float, true, false, true, 0, 1>);
  static FdmRegistrar f3(fdm_key(3, 6, 3, 2), &launch_fdm<3, 6, 3, float, double, double>);
  static OpRegistrar r4(op_key(3, 7, 1, 0, 1, 0, 0, 0, 1), &launch_op<3, 7, 1, double, true, false, false, 0, 1>);
  static OpRegistrar r5(op_key(2, 7, 4, 0, 0, 1, 0, 1, 0), &launch_op<2, 7, 4, double, false, true, false, 1, 0>);
  static BatchRegistrar s6(batch_key(3, 6), &launch_batch<3, 6>);
  static OpRegistrar r7(op_key(3, 4, 4, 1, 1, 1, 0, 1, 0), &launch_op<3, 4, 4, float, true, true, false, 1, 0>);
  static OpRegistrar r8(op_key(3, 4, 3, 0, 0, 1, 0, 1, 0), &launch_op<3, 4, 3, double, false, true, false, 1, 0>);
  static FdmRegistrar f9(fdm_key(2, 6, 3, 0), &launch_fdm<2, 6, 3, double, double, double>);
  static OpRegistrar r10(op_key(2, 8, 1, 1, 0, 1, 0, 1, 0), &launch_op<2, 8, 1, float, false, true, false, 1, 0>);
  static FdmRegistrar f11(fdm_key(2, 5, 4, 1), &launch_fdm<2, 5, 4, float, float, float>);
  static OpRegistrar r12(op_key(2, 7, 1, 0, 1, 0, 0, 0, 0), &launch_op<2, 7, 1, double, true, false, false, 0, 0>);
  static OpRegistrar r13(op_key(3, 4, 1, 0, 0, 1, 0, 1, 0), &launch_op<3, 4, 1, double, false, true, false, 1, 0>);
  static OpRegistrar r14(op_key(2, 5, 1, 1, 1, 1, 1, 1, 0), &launch_op<2, 5, 1, float, true, true, true, 1, 0>);
  static BurgersRegistrar g15(burgers_key(2, 4, 0, 0), &launch_burgers<2, 4, 6, false, 0>);
  static OpRegistrar r16(op_key(2, 4, 2, 0, 0, 1, 0, 1, 0), &launch_op<2, 4, 2, double, false, true, false, 1, 0>);
  static FdmRegistrar f17(fdm_key(2, 5, 1, 1), &launch_fdm<2, 5, 1, float, float, float>);
  static FdmRegistrar f18(fdm_key(2, 3, 3, 1), &launch_fdm<2, 3, 3, float, float, float>);
  static FdmRegistrar f19(fdm_key(2, 4, 1, 0), &launch_fdm<2, 4, 1, double, double, double>);
  static FdmRegistrar f20(fdm_key(3, 2, 3, 1), &launch_fdm<3, 2, 3, float, float, float>);
  static OpRegistrar r21(op_key(2, 3, 1, 0, 1, 0, 0, 0, 0), &launch_op<2, 3, 1, double, true, false, false, 0, 0>);
  static OpRegistrar r22(op_key(3, 2, 1, 0, 1, 0, 0, 0, 0), &launch_op<3, 2, 1, double, true, false, false, 0, 0>);
  static OpRegistrar r23(op_key(2, 2, 1, 0, 0, 0, 0, 0, 0), &launch_op<2, 2, 1, double, false, false, false, 0, 0>);
} // namespace dgops
