#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 3, 1), &launch_fdm<3, 8, 3, float, float, float>);
  static OpRegistrar r1(op_key(3, 8, 1, 1, 1, 0, 1, 0, 1), &launch_op<3, 8, 1, float, 1, false, true, 0, 1>);
  static OpRegistrar r2(op_key(3, 6, 4, 0, 2, 1, 0, 2, 0), &launch_op<3, 6, 4, double, 2, true, false, 2, 0>);
  static OpRegistrar r3(op_key(3, 7, 1, 1, 2, 0, 1, 0, 1), &launch_op<3, 7, 1, float, 2, false, true, 0, 1>);
  static OpRegistrar r4(op_key(3, 6, 3, 1, 2, 1, 0, 1, 0), &launch_op<3, 6, 3, float, 2, true, false, 1, 0>);
  static OpRegistrar r5(op_key(3, 6, 1, 1, 1, 1, 1, 1, 0), &launch_op<3, 6, 1, float, 1, true, true, 1, 0>);
  static OpRegistrar r6(op_key(3, 7, 1, 0, 0, 1, 0, 1, 0), &launch_op<3, 7, 1, double, 0, true, false, 1, 0>);
  static FdmRegistrar f7(fdm_key(2, 8, 4, 0), &launch_fdm<2, 8, 4, double, double, double>);
  static OpRegistrar r8(op_key(2, 7, 4, 1, 2, 1, 0, 2, 0), &launch_op<2, 7, 4, float, 2, true, false, 2, 0>);
  static BurgersRegistrar g9(burgers_key(3, 4, 1, 1), &launch_burgers<3, 4, 6, true, 1>);
  static OpRegistrar r10(op_key(3, 5, 1, 0, 1, 1, 1, 1, 0), &launch_op<3, 5, 1, double, 1, true, true, 1, 0>);
  static OpRegistrar r11(op_key(3, 4, 4, 0, 0, 1, 0, 1, 0), &launch_op<3, 4, 4, double, 0, true, false, 1, 0>);
  static FdmRegistrar f12(fdm_key(3, 4, 4, 1), &launch_fdm<3, 4, 4, float, float, float>);
  static OpRegistrar r13(op_key(3, 4, 3, 1, 2, 1, 0, 2, 0), &launch_op<3, 4, 3, float, 2, true, false, 2, 0>);
  static FdmRegistrar f14(fdm_key(2, 7, 2, 0), &launch_fdm<2, 7, 2, double, double, double>);
  static OpRegistrar r15(op_key(3, 5, 1, 0, 0, 1, 0, 2, 0), &launch_op<3, 5, 1, double, 0, true, false, 2, 0>);
  static OpRegistrar r16(op_key(2, 8, 1, 1, 0, 0, 0, 0, 0), &launch_op<2, 8, 1, float, 0, false, false, 0, 0>);
  static OpRegistrar r17(op_key(3, 4, 1, 1, 2, 0, 1, 0, 1), &launch_op<3, 4, 1, float, 2, false, true, 0, 1>);
  static OpRegistrar r18(op_key(2, 6, 2, 1, 2, 1, 0, 1, 0), &launch_op<2, 6, 2, float, 2, true, false, 1, 0>);
  static OpRegistrar r19(op_key(2, 5, 3, 0, 0, 1, 0, 2, 0), &launch_op<2, 5, 3, double, 0, true, false, 2, 0>);
  static MassRegistrar m20(mass_key(2, 7, 0), &launch_mass<2, 7, double>);
  static OpRegistrar r21(op_key(3, 4, 1, 0, 0, 0, 0, 0, 1), &launch_op<3, 4, 1, double, 0, false, false, 0, 1>);
  static FdmRegistrar f22(fdm_key(3, 4, 1, 0), &launch_fdm<3, 4, 1, double, double, double>);
  static OpRegistrar r23(op_key(2, 5, 1, 0, 1, 0, 1, 0, 0), &launch_op<2, 5, 1, double, 1, false, true, 0, 0>);
  static FdmRegistrar f24(fdm_key(3, 3, 3, 1), &launch_fdm<3, 3, 3, float, float, float>);
  static MassRegistrar m25(mass_key(2, 6, 1), &launch_mass<2, 6, float>);
  static OpRegistrar r26(op_key(3, 3, 2, 1, 0, 1, 0, 1, 0), &launch_op<3, 3, 2, float, 0, true, false, 1, 0>);
  static OpRegistrar r27(op_key(2, 4, 2, 1, 2, 1, 0, 1, 0), &launch_op<2, 4, 2, float, 2, true, false, 1, 0>);
  static OpRegistrar r28(op_key(2, 5, 1, 1, 0, 0, 0, 0, 0), &launch_op<2, 5, 1, float, 0, false, false, 0, 0>);
  static CellGmresRegistrar q29(cell_gmres_key(2, 5, 4, 1), &launch_cell_gmres<2, 5, 4, float>);
  static OpRegistrar r30(op_key(2, 3, 3, 1, 0, 1, 0, 1, 0), &launch_op<2, 3, 3, float, 0, true, false, 1, 0>);
  static CellGmresRegistrar q31(cell_gmres_key(3, 3, 1, 0), &launch_cell_gmres<3, 3, 1, double>);
  static OpRegistrar r32(op_key(2, 4, 1, 0, 2, 1, 0, 2, 0), &launch_op<2, 4, 1, double, 2, true, false, 2, 0>);
  static CellGmresRegistrar q33(cell_gmres_key(2, 4, 4, 1), &launch_cell_gmres<2, 4, 4, float>);
  static OpRegistrar r34(op_key(3, 2, 3, 0, 0, 1, 0, 1, 0), &launch_op<3, 2, 3, double, 0, true, false, 1, 0>);
  static OpRegistrar r35(op_key(3, 2, 2, 0, 2, 1, 0, 2, 0), &launch_op<3, 2, 2, double, 2, true, false, 2, 0>);
  static OpRegistrar r36(op_key(2, 3, 1, 0, 2, 0, 0, 0, 1), &launch_op<2, 3, 1, double, 2, false, false, 0, 1>);
  static CellGmresRegistrar q37(cell_gmres_key(2, 3, 4, 0), &launch_cell_gmres<2, 3, 4, double>);
  static OpRegistrar r38(op_key(3, 2, 1, 1, 0, 1, 0, 1, 0), &launch_op<3, 2, 1, float, 0, true, false, 1, 0>);
  static OpRegistrar r39(op_key(2, 2, 1, 1, 2, 0, 1, 0, 0), &launch_op<2, 2, 1, float, 2, false, true, 0, 0>);
  static OpRegistrar r40(op_key(2, 2, 1, 0, 2, 0, 0, 0, 0), &launch_op<2, 2, 1, double, 2, false, false, 0, 0>);
  static CellGmresRegistrar q41(cell_gmres_key(2, 2, 2, 1), &launch_cell_gmres<2, 2, 2, float>);
} // namespace dgops
