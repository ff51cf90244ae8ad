#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 8, 1, 0, 2, 0, 1, 0, 1), &launch_op<3, 8, 1, double, 2, false, true, 0, 1>);
  static FdmRegistrar f1(fdm_key(3, 6, 4, 2), &launch_fdm<3, 6, 4, float, double, double>);
  static OpRegistrar r2(op_key(3, 8, 1, 1, 0, 0, 0, 0, 0), &launch_op<3, 8, 1, float, 0, false, false, 0, 0>);
  static OpRegistrar r3(op_key(3, 6, 1, 0, 1, 0, 1, 0, 1), &launch_op<3, 6, 1, double, 1, false, true, 0, 1>);
  static OpRegistrar r4(op_key(3, 7, 1, 0, 2, 0, 0, 0, 1), &launch_op<3, 7, 1, double, 2, false, false, 0, 1>);
  static FdmRegistrar f5(fdm_key(2, 8, 3, 1), &launch_fdm<2, 8, 3, float, float, float>);
  static BurgersRegistrar g6(burgers_key(3, 4, 1, 0), &launch_burgers<3, 4, 6, true, 0>);
  static OpRegistrar r7(op_key(2, 7, 3, 0, 0, 1, 0, 1, 0), &launch_op<2, 7, 3, double, 0, true, false, 1, 0>);
  static FdmRegistrar f8(fdm_key(2, 8, 2, 1), &launch_fdm<2, 8, 2, float, float, float>);
  static OpRegistrar r9(op_key(2, 7, 2, 0, 2, 1, 0, 1, 0), &launch_op<2, 7, 2, double, 2, true, false, 1, 0>);
  static FdmRegistrar f10(fdm_key(2, 6, 3, 2), &launch_fdm<2, 6, 3, float, double, double>);
  static OpRegistrar r11(op_key(2, 8, 1, 1, 0, 1, 0, 1, 0), &launch_op<2, 8, 1, float, 0, true, false, 1, 0>);
  static FdmRegistrar f12(fdm_key(3, 4, 2, 2), &launch_fdm<3, 4, 2, float, double, double>);
  static OpRegistrar r13(op_key(2, 6, 1, 1, 2, 0, 1, 0, 1), &launch_op<2, 6, 1, float, 2, false, true, 0, 1>);
  static FdmRegistrar f14(fdm_key(2, 7, 1, 2), &launch_fdm<2, 7, 1, float, double, double>);
  static FdmRegistrar f15(fdm_key(2, 4, 4, 0), &launch_fdm<2, 4, 4, double, double, double>);
  static OpRegistrar r16(op_key(2, 5, 1, 1, 2, 1, 1, 1, 0), &launch_op<2, 5, 1, float, 2, true, true, 1, 0>);
  static BurgersRegistrar g17(burgers_key(2, 4, 0, 0), &launch_burgers<2, 4, 6, false, 0>);
  static OpRegistrar r18(op_key(3, 3, 1, 1, 1, 0, 1, 0, 1), &launch_op<3, 3, 1, float, 1, false, true, 0, 1>);
  static FdmRegistrar f19(fdm_key(2, 4, 2, 0), &launch_fdm<2, 4, 2, double, double, double>);
  static FdmRegistrar f20(fdm_key(2, 3, 4, 2), &launch_fdm<2, 3, 4, float, double, double>);
  static BurgersRegistrar g21(burgers_key(3, 3, 1, 1), &launch_burgers<3, 3, 3, true, 1>);
  static MassRegistrar m22(mass_key(2, 4, 0), &launch_mass<2, 4, double>);
  static OpRegistrar r23(op_key(3, 2, 3, 1, 2, 1, 0, 1, 0), &launch_op<3, 2, 3, float, 2, true, false, 1, 0>);
  static FdmRegistrar f24(fdm_key(2, 2, 4, 0), &launch_fdm<2, 2, 4, double, double, double>);
  static MassRegistrar m25(mass_key(2, 3, 1), &launch_mass<2, 3, float>);
  static OpRegistrar r26(op_key(2, 2, 1, 0, 1, 1, 1, 1, 0), &launch_op<2, 2, 1, double, 1, true, true, 1, 0>);
  static OpRegistrar r27(op_key(2, 2, 1, 1, 2, 0, 0, 0, 0), &launch_op<2, 2, 1, float, 2, false, false, 0, 0>);
} // namespace dgops
