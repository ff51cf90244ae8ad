#include "../registry.hpp"

namespace dgops
{
  static BurgersRegistrar g0(burgers_key(3, 5, 0, 1), &launch_burgers<3, 5, 10, false, 1>);
  static FdmRegistrar f1(fdm_key(3, 8, 2, 2), &launch_fdm<3, 8, 2, float, double, double>);
  static OpRegistrar r2(op_key(3, 7, 1, 0, 2, 1, 1, 1, 0), &launch_op<3, 7, 1, double, 2, true, true, 1, 0>);
  static MassRegistrar m3(mass_key(3, 8, 0), &launch_mass<3, 8, double>);
  static OpRegistrar r4(op_key(3, 6, 1, 1, 2, 1, 1, 1, 0), &launch_op<3, 6, 1, float, 2, true, true, 1, 0>);
  static BatchRegistrar s5(batch_key(3, 7), &launch_batch<3, 7>);
  static OpRegistrar r6(op_key(3, 6, 1, 0, 0, 0, 0, 0, 1), &launch_op<3, 6, 1, double, 0, false, false, 0, 1>);
  static OpRegistrar r7(op_key(3, 5, 1, 0, 1, 0, 1, 0, 1), &launch_op<3, 5, 1, double, 1, false, true, 0, 1>);
  static OpRegistrar r8(op_key(3, 4, 4, 1, 2, 1, 0, 1, 0), &launch_op<3, 4, 4, float, 2, true, false, 1, 0>);
  static OpRegistrar r9(op_key(2, 6, 4, 0, 2, 1, 0, 1, 0), &launch_op<2, 6, 4, double, 2, true, false, 1, 0>);
  static OpRegistrar r10(op_key(2, 7, 1, 1, 1, 0, 1, 0, 1), &launch_op<2, 7, 1, float, 1, false, true, 0, 1>);
  static OpRegistrar r11(op_key(3, 5, 1, 1, 2, 1, 0, 1, 0), &launch_op<3, 5, 1, float, 2, true, false, 1, 0>);
  static OpRegistrar r12(op_key(3, 4, 1, 0, 2, 0, 1, 0, 0), &launch_op<3, 4, 1, double, 2, false, true, 0, 0>);
  static OpRegistrar r13(op_key(2, 6, 2, 0, 0, 1, 0, 1, 0), &launch_op<2, 6, 2, double, 0, true, false, 1, 0>);
  static FdmRegistrar f14(fdm_key(2, 5, 3, 1), &launch_fdm<2, 5, 3, float, float, float>);
  static OpRegistrar r15(op_key(3, 4, 1, 0, 0, 0, 0, 0, 0), &launch_op<3, 4, 1, double, 0, false, false, 0, 0>);
  static OpRegistrar r16(op_key(2, 5, 2, 0, 2, 1, 0, 1, 0), &launch_op<2, 5, 2, double, 2, true, false, 1, 0>);
  static FdmRegistrar f17(fdm_key(3, 3, 3, 2), &launch_fdm<3, 3, 3, float, double, double>);
  static FdmRegistrar f18(fdm_key(2, 4, 3, 0), &launch_fdm<2, 4, 3, double, double, double>);
  static OpRegistrar r19(op_key(2, 4, 2, 1, 0, 1, 0, 1, 0), &launch_op<2, 4, 2, float, 0, true, false, 1, 0>);
  static OpRegistrar r20(op_key(2, 5, 1, 1, 0, 1, 0, 1, 0), &launch_op<2, 5, 1, float, 0, true, false, 1, 0>);
  static OpRegistrar r21(op_key(3, 3, 1, 1, 2, 1, 0, 1, 0), &launch_op<3, 3, 1, float, 2, true, false, 1, 0>);
  static OpRegistrar r22(op_key(2, 4, 1, 0, 0, 1, 0, 1, 0), &launch_op<2, 4, 1, double, 0, true, false, 1, 0>);
  static OpRegistrar r23(op_key(2, 3, 1, 0, 2, 0, 1, 0, 1), &launch_op<2, 3, 1, double, 2, false, true, 0, 1>);
  static OpRegistrar r24(op_key(3, 2, 2, 1, 0, 1, 0, 1, 0), &launch_op<3, 2, 2, float, 0, true, false, 1, 0>);
  static OpRegistrar r25(op_key(2, 3, 1, 1, 0, 0, 0, 0, 0), &launch_op<2, 3, 1, float, 0, false, false, 0, 0>);
  static OpRegistrar r26(op_key(3, 2, 1, 0, 2, 0, 0, 0, 0), &launch_op<3, 2, 1, double, 2, false, false, 0, 0>);
  static FdmRegistrar f27(fdm_key(2, 2, 2, 0), &launch_fdm<2, 2, 2, double, double, double>);
  static BurgersRegistrar g28(burgers_key(2, 2, 0, 0), &launch_burgers<2, 2, 1, false, 0>);
} // namespace dgops
