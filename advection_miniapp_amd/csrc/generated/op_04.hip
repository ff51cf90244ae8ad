#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 3, 1), &launch_fdm<3, 8, 3, float, float, float>);
  static OpRegistrar r1(op_key(3, 8, 1, 1, 2, 1, 1, 1, 0), &launch_op<3, 8, 1, float, 2, true, true, 1, 0>);
  static OpRegistrar r2(op_key(3, 7, 1, 0, 1, 1, 1, 1, 0), &launch_op<3, 7, 1, double, 1, true, true, 1, 0>);
  static FdmRegistrar f3(fdm_key(3, 8, 1, 1), &launch_fdm<3, 8, 1, float, float, float>);
  static OpRegistrar r4(op_key(3, 6, 1, 1, 1, 1, 1, 1, 0), &launch_op<3, 6, 1, float, 1, true, true, 1, 0>);
  static FdmRegistrar f5(fdm_key(3, 7, 1, 0), &launch_fdm<3, 7, 1, double, double, double>);
  static FdmRegistrar f6(fdm_key(2, 7, 4, 1), &launch_fdm<2, 7, 4, float, float, float>);
  static OpRegistrar r7(op_key(3, 5, 2, 1, 0, 1, 0, 1, 0), &launch_op<3, 5, 2, float, 0, true, false, 1, 0>);
  static OpRegistrar r8(op_key(3, 4, 4, 0, 0, 1, 0, 1, 0), &launch_op<3, 4, 4, double, 0, true, false, 1, 0>);
  static BurgersRegistrar g9(burgers_key(2, 5, 1, 0), &launch_burgers<2, 5, 10, true, 0>);
  static OpRegistrar r10(op_key(2, 7, 1, 0, 2, 0, 1, 0, 1), &launch_op<2, 7, 1, double, 2, false, true, 0, 1>);
  static OpRegistrar r11(op_key(3, 5, 1, 0, 0, 1, 0, 1, 0), &launch_op<3, 5, 1, double, 0, true, false, 1, 0>);
  static OpRegistrar r12(op_key(3, 4, 1, 0, 1, 0, 1, 0, 0), &launch_op<3, 4, 1, double, 1, false, true, 0, 0>);
  static FdmRegistrar f13(fdm_key(2, 5, 4, 0), &launch_fdm<2, 5, 4, double, double, double>);
  static OpRegistrar r14(op_key(2, 5, 3, 1, 0, 1, 0, 1, 0), &launch_op<2, 5, 3, float, 0, true, false, 1, 0>);
  static FdmRegistrar f15(fdm_key(3, 3, 4, 0), &launch_fdm<3, 3, 4, double, double, double>);
  static MassRegistrar m16(mass_key(3, 4, 0), &launch_mass<3, 4, double>);
  static OpRegistrar r17(op_key(3, 3, 3, 1, 2, 1, 0, 1, 0), &launch_op<3, 3, 3, float, 2, true, false, 1, 0>);
  static OpRegistrar r18(op_key(2, 4, 3, 0, 2, 1, 0, 1, 0), &launch_op<2, 4, 3, double, 2, true, false, 1, 0>);
  static FdmRegistrar f19(fdm_key(3, 3, 2, 2), &launch_fdm<3, 3, 2, float, double, double>);
  static OpRegistrar r20(op_key(2, 5, 1, 1, 0, 0, 0, 0, 0), &launch_op<2, 5, 1, float, 0, false, false, 0, 0>);
  static OpRegistrar r21(op_key(3, 3, 1, 0, 0, 1, 0, 1, 0), &launch_op<3, 3, 1, double, 0, true, false, 1, 0>);
  static OpRegistrar r22(op_key(2, 4, 1, 0, 2, 0, 0, 0, 1), &launch_op<2, 4, 1, double, 2, false, false, 0, 1>);
  static OpRegistrar r23(op_key(2, 3, 1, 0, 1, 0, 1, 0, 1), &launch_op<2, 3, 1, double, 1, false, true, 0, 1>);
  static OpRegistrar r24(op_key(2, 2, 4, 1, 2, 1, 0, 1, 0), &launch_op<2, 2, 4, float, 2, true, false, 1, 0>);
  static OpRegistrar r25(op_key(2, 3, 1, 0, 0, 0, 0, 0, 1), &launch_op<2, 3, 1, double, 0, false, false, 0, 1>);
  static FdmRegistrar f26(fdm_key(2, 2, 3, 2), &launch_fdm<2, 2, 3, float, double, double>);
  static OpRegistrar r27(op_key(2, 2, 1, 1, 2, 0, 1, 0, 0), &launch_op<2, 2, 1, float, 2, false, true, 0, 0>);
  static FdmRegistrar f28(fdm_key(2, 2, 1, 2), &launch_fdm<2, 2, 1, float, double, double>);
} // namespace dgops
