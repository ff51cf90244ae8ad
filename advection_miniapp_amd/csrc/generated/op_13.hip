#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 7, 4, 1, 2, 1, 0, 1, 0), &launch_op<3, 7, 4, float, 2, true, false, 1, 0>);
  static FdmRegistrar f1(fdm_key(3, 7, 3, 1), &launch_fdm<3, 7, 3, float, float, float>);
  static OpRegistrar r2(op_key(3, 7, 1, 1, 2, 1, 1, 1, 0), &launch_op<3, 7, 1, float, 2, true, true, 1, 0>);
  static FdmRegistrar f3(fdm_key(3, 6, 3, 0), &launch_fdm<3, 6, 3, double, double, double>);
  static OpRegistrar r4(op_key(3, 5, 4, 1, 0, 1, 0, 1, 0), &launch_op<3, 5, 4, float, 0, true, false, 1, 0>);
  static OpRegistrar r5(op_key(3, 5, 3, 0, 0, 1, 0, 1, 0), &launch_op<3, 5, 3, double, 0, true, false, 1, 0>);
  static OpRegistrar r6(op_key(3, 6, 1, 1, 0, 1, 0, 1, 0), &launch_op<3, 6, 1, float, 0, true, false, 1, 0>);
  static OpRegistrar r7(op_key(3, 5, 1, 1, 1, 0, 1, 0, 1), &launch_op<3, 5, 1, float, 1, false, true, 0, 1>);
  static OpRegistrar r8(op_key(2, 8, 1, 0, 2, 1, 1, 1, 0), &launch_op<2, 8, 1, double, 2, true, true, 1, 0>);
  static OpRegistrar r9(op_key(3, 4, 3, 0, 0, 1, 0, 1, 0), &launch_op<3, 4, 3, double, 0, true, false, 1, 0>);
  static FdmRegistrar f10(fdm_key(2, 7, 2, 1), &launch_fdm<2, 7, 2, float, float, float>);
  static MassRegistrar m11(mass_key(3, 5, 1), &launch_mass<3, 5, float>);
  static OpRegistrar r12(op_key(3, 4, 1, 1, 2, 0, 1, 0, 0), &launch_op<3, 4, 1, float, 2, false, true, 0, 0>);
  static OpRegistrar r13(op_key(2, 6, 1, 0, 1, 1, 1, 1, 0), &launch_op<2, 6, 1, double, 1, true, true, 1, 0>);
  static OpRegistrar r14(op_key(2, 7, 1, 1, 0, 0, 0, 0, 0), &launch_op<2, 7, 1, float, 0, false, false, 0, 0>);
  static OpRegistrar r15(op_key(3, 4, 1, 0, 0, 1, 0, 1, 0), &launch_op<3, 4, 1, double, 0, true, false, 1, 0>);
  static OpRegistrar r16(op_key(2, 5, 1, 0, 2, 0, 1, 0, 0), &launch_op<2, 5, 1, double, 2, false, true, 0, 0>);
  static OpRegistrar r17(op_key(2, 6, 1, 1, 2, 0, 0, 0, 0), &launch_op<2, 6, 1, float, 2, false, false, 0, 0>);
  static OpRegistrar r18(op_key(3, 3, 2, 1, 2, 1, 0, 1, 0), &launch_op<3, 3, 2, float, 2, true, false, 1, 0>);
  static OpRegistrar r19(op_key(2, 4, 1, 0, 2, 0, 1, 0, 1), &launch_op<2, 4, 1, double, 2, false, true, 0, 1>);
  static MassRegistrar m20(mass_key(2, 5, 0), &launch_mass<2, 5, double>);
  static FdmRegistrar f21(fdm_key(2, 3, 3, 1), &launch_fdm<2, 3, 3, float, float, float>);
  static OpRegistrar r22(op_key(3, 2, 4, 1, 0, 1, 0, 1, 0), &launch_op<3, 2, 4, float, 0, true, false, 1, 0>);
  static OpRegistrar r23(op_key(2, 3, 1, 1, 2, 0, 1, 0, 1), &launch_op<2, 3, 1, float, 2, false, true, 0, 1>);
  static OpRegistrar r24(op_key(3, 2, 1, 0, 2, 0, 1, 0, 1), &launch_op<3, 2, 1, double, 2, false, true, 0, 1>);
  static FdmRegistrar f25(fdm_key(2, 3, 1, 0), &launch_fdm<2, 3, 1, double, double, double>);
  static OpRegistrar r26(op_key(3, 2, 1, 0, 2, 1, 0, 1, 0), &launch_op<3, 2, 1, double, 2, true, false, 1, 0>);
  static BatchRegistrar s27(batch_key(3, 2), &launch_batch<3, 2>);
  static BurgersRegistrar g28(burgers_key(3, 2, 1, 0), &launch_burgers<3, 2, 1, true, 0>);
} // namespace dgops
