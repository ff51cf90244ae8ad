#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 4, 1), &launch_fdm<3, 8, 4, float, float, float>);
  static OpRegistrar r1(op_key(3, 8, 1, 1, 1, 1, 1, 1, 0), &launch_op<3, 8, 1, float, 1, true, true, 1, 0>);
  static OpRegistrar r2(op_key(3, 7, 2, 1, 2, 1, 0, 1, 0), &launch_op<3, 7, 2, float, 2, true, false, 1, 0>);
  static OpRegistrar r3(op_key(3, 8, 1, 0, 2, 1, 0, 1, 0), &launch_op<3, 8, 1, double, 2, true, false, 1, 0>);
  static OpRegistrar r4(op_key(3, 6, 1, 0, 2, 1, 1, 1, 0), &launch_op<3, 6, 1, double, 2, true, true, 1, 0>);
  static OpRegistrar r5(op_key(3, 7, 1, 1, 0, 1, 0, 1, 0), &launch_op<3, 7, 1, float, 0, true, false, 1, 0>);
  static OpRegistrar r6(op_key(2, 7, 4, 1, 0, 1, 0, 1, 0), &launch_op<2, 7, 4, float, 0, true, false, 1, 0>);
  static MassRegistrar m7(mass_key(3, 6, 1), &launch_mass<3, 6, float>);
  static FdmRegistrar f8(fdm_key(2, 7, 3, 0), &launch_fdm<2, 7, 3, double, double, double>);
  static FdmRegistrar f9(fdm_key(3, 4, 4, 2), &launch_fdm<3, 4, 4, float, double, double>);
  static OpRegistrar r10(op_key(2, 7, 1, 0, 1, 0, 1, 0, 1), &launch_op<2, 7, 1, double, 1, false, true, 0, 1>);
  static OpRegistrar r11(op_key(3, 5, 1, 0, 2, 0, 0, 0, 1), &launch_op<3, 5, 1, double, 2, false, false, 0, 1>);
  static OpRegistrar r12(op_key(3, 4, 2, 0, 2, 1, 0, 1, 0), &launch_op<3, 4, 2, double, 2, true, false, 1, 0>);
  static OpRegistrar r13(op_key(2, 5, 4, 0, 2, 1, 0, 1, 0), &launch_op<2, 5, 4, double, 2, true, false, 1, 0>);
  static FdmRegistrar f14(fdm_key(2, 6, 2, 2), &launch_fdm<2, 6, 2, float, double, double>);
  static OpRegistrar r15(op_key(3, 3, 4, 0, 2, 1, 0, 1, 0), &launch_op<3, 3, 4, double, 2, true, false, 1, 0>);
  static FdmRegistrar f16(fdm_key(3, 4, 1, 1), &launch_fdm<3, 4, 1, float, float, float>);
  static OpRegistrar r17(op_key(3, 3, 3, 0, 0, 1, 0, 1, 0), &launch_op<3, 3, 3, double, 0, true, false, 1, 0>);
  static MassRegistrar m18(mass_key(2, 6, 0), &launch_mass<2, 6, double>);
  static OpRegistrar r19(op_key(3, 3, 1, 1, 2, 1, 1, 1, 0), &launch_op<3, 3, 1, float, 2, true, true, 1, 0>);
  static OpRegistrar r20(op_key(2, 5, 1, 0, 0, 0, 0, 0, 1), &launch_op<2, 5, 1, double, 0, false, false, 0, 1>);
  static OpRegistrar r21(op_key(3, 3, 1, 0, 2, 0, 0, 0, 1), &launch_op<3, 3, 1, double, 2, false, false, 0, 1>);
  static OpRegistrar r22(op_key(2, 4, 1, 0, 0, 0, 0, 0, 0), &launch_op<2, 4, 1, double, 0, false, false, 0, 0>);
  static OpRegistrar r23(op_key(2, 3, 2, 1, 0, 1, 0, 1, 0), &launch_op<2, 3, 2, float, 0, true, false, 1, 0>);
  static OpRegistrar r24(op_key(2, 2, 4, 0, 0, 1, 0, 1, 0), &launch_op<2, 2, 4, double, 0, true, false, 1, 0>);
  static FdmRegistrar f25(fdm_key(3, 2, 2, 1), &launch_fdm<3, 2, 2, float, float, float>);
  static OpRegistrar r26(op_key(2, 2, 3, 1, 2, 1, 0, 1, 0), &launch_op<2, 2, 3, float, 2, true, false, 1, 0>);
  static OpRegistrar r27(op_key(2, 2, 1, 1, 1, 0, 1, 0, 0), &launch_op<2, 2, 1, float, 1, false, true, 0, 0>);
  static OpRegistrar r28(op_key(2, 2, 1, 1, 2, 1, 0, 1, 0), &launch_op<2, 2, 1, float, 2, true, false, 1, 0>);
} // namespace dgops
