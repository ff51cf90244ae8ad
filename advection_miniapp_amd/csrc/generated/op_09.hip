#include "../registry.hpp"

namespace dgops
{
  static BurgersRegistrar g0(burgers_key(3, 5, 1, 1), &launch_burgers<3, 5, 10, true, 1>);
  static FdmRegistrar f1(fdm_key(3, 8, 2, 0), &launch_fdm<3, 8, 2, double, double, double>);
  static FdmRegistrar f2(fdm_key(3, 6, 4, 2), &launch_fdm<3, 6, 4, float, double, double>);
  static OpRegistrar r3(op_key(3, 8, 1, 0, 0, 0, 0, 0, 0), &launch_op<3, 8, 1, double, 0, false, false, 0, 0>);
  static FdmRegistrar f4(fdm_key(3, 6, 3, 0), &launch_fdm<3, 6, 3, double, double, double>);
  static FdmRegistrar f5(fdm_key(3, 6, 2, 1), &launch_fdm<3, 6, 2, float, float, float>);
  static OpRegistrar r6(op_key(3, 7, 1, 0, 2, 1, 0, 2, 0), &launch_op<3, 7, 1, double, 2, true, false, 2, 0>);
  static OpRegistrar r7(op_key(3, 5, 3, 1, 0, 1, 0, 1, 0), &launch_op<3, 5, 3, float, 0, true, false, 1, 0>);
  static OpRegistrar r8(op_key(3, 6, 1, 0, 0, 0, 0, 0, 1), &launch_op<3, 6, 1, double, 0, false, false, 0, 1>);
  static CellGmresRegistrar q9(cell_gmres_key(3, 6, 1, 0), &launch_cell_gmres<3, 6, 1, double>);
  static OpRegistrar r10(op_key(3, 5, 1, 1, 1, 0, 1, 0, 1), &launch_op<3, 5, 1, float, 1, false, true, 0, 1>);
  static OpRegistrar r11(op_key(3, 4, 4, 0, 2, 1, 0, 2, 0), &launch_op<3, 4, 4, double, 2, true, false, 2, 0>);
  static BurgersRegistrar g12(burgers_key(2, 5, 1, 1), &launch_burgers<2, 5, 10, true, 1>);
  static OpRegistrar r13(op_key(2, 7, 2, 0, 2, 1, 0, 1, 0), &launch_op<2, 7, 2, double, 2, true, false, 1, 0>);
  static OpRegistrar r14(op_key(2, 6, 3, 1, 0, 1, 0, 1, 0), &launch_op<2, 6, 3, float, 0, true, false, 1, 0>);
  static FdmRegistrar f15(fdm_key(3, 5, 1, 1), &launch_fdm<3, 5, 1, float, float, float>);
  static OpRegistrar r16(op_key(2, 8, 1, 1, 2, 1, 0, 1, 0), &launch_op<2, 8, 1, float, 2, true, false, 1, 0>);
  static FdmRegistrar f17(fdm_key(3, 4, 2, 0), &launch_fdm<3, 4, 2, double, double, double>);
  static OpRegistrar r18(op_key(2, 6, 1, 0, 1, 0, 1, 0, 0), &launch_op<2, 6, 1, double, 1, false, true, 0, 0>);
  static FdmRegistrar f19(fdm_key(2, 5, 3, 1), &launch_fdm<2, 5, 3, float, float, float>);
  static CellGmresRegistrar q20(cell_gmres_key(2, 7, 3, 0), &launch_cell_gmres<2, 7, 3, double>);
  static OpRegistrar r21(op_key(3, 4, 1, 0, 0, 1, 0, 1, 0), &launch_op<3, 4, 1, double, 0, true, false, 1, 0>);
  static MassRegistrar m22(mass_key(3, 4, 1), &launch_mass<3, 4, float>);
  static OpRegistrar r23(op_key(2, 5, 1, 0, 2, 1, 1, 1, 0), &launch_op<2, 5, 1, double, 2, true, true, 1, 0>);
  static OpRegistrar r24(op_key(2, 6, 1, 0, 2, 0, 0, 0, 1), &launch_op<2, 6, 1, double, 2, false, false, 0, 1>);
  static CellGmresRegistrar q25(cell_gmres_key(2, 6, 4, 0), &launch_cell_gmres<2, 6, 4, double>);
  static OpRegistrar r26(op_key(3, 3, 2, 1, 2, 1, 0, 2, 0), &launch_op<3, 3, 2, float, 2, true, false, 2, 0>);
  static OpRegistrar r27(op_key(2, 4, 1, 0, 1, 0, 1, 0, 0), &launch_op<2, 4, 1, double, 1, false, true, 0, 0>);
  static OpRegistrar r28(op_key(2, 5, 1, 1, 2, 1, 0, 1, 0), &launch_op<2, 5, 1, float, 2, true, false, 1, 0>);
  static OpRegistrar r29(op_key(2, 3, 4, 0, 0, 1, 0, 2, 0), &launch_op<2, 3, 4, double, 0, true, false, 2, 0>);
  static OpRegistrar r30(op_key(2, 3, 3, 1, 2, 1, 0, 2, 0), &launch_op<2, 3, 3, float, 2, true, false, 2, 0>);
  static CellGmresRegistrar q31(cell_gmres_key(3, 3, 1, 1), &launch_cell_gmres<3, 3, 1, float>);
  static OpRegistrar r32(op_key(3, 2, 4, 1, 0, 1, 0, 2, 0), &launch_op<3, 2, 4, float, 0, true, false, 2, 0>);
  static OpRegistrar r33(op_key(2, 3, 2, 0, 0, 1, 0, 2, 0), &launch_op<2, 3, 2, double, 0, true, false, 2, 0>);
  static OpRegistrar r34(op_key(3, 2, 3, 0, 2, 1, 0, 2, 0), &launch_op<3, 2, 3, double, 2, true, false, 2, 0>);
  static OpRegistrar r35(op_key(3, 2, 1, 0, 1, 1, 1, 1, 0), &launch_op<3, 2, 1, double, 1, true, true, 1, 0>);
  static OpRegistrar r36(op_key(2, 3, 1, 0, 2, 1, 0, 1, 0), &launch_op<2, 3, 1, double, 2, true, false, 1, 0>);
  static CellGmresRegistrar q37(cell_gmres_key(2, 3, 4, 1), &launch_cell_gmres<2, 3, 4, float>);
  static OpRegistrar r38(op_key(2, 2, 2, 1, 0, 1, 0, 1, 0), &launch_op<2, 2, 2, float, 0, true, false, 1, 0>);
  static FdmRegistrar f39(fdm_key(2, 2, 2, 2), &launch_fdm<2, 2, 2, float, double, double>);
  static OpRegistrar r40(op_key(2, 2, 1, 1, 0, 1, 0, 1, 0), &launch_op<2, 2, 1, float, 0, true, false, 1, 0>);
  static BurgersRegistrar g41(burgers_key(2, 2, 1, 0), &launch_burgers<2, 2, 1, true, 0>);
} // namespace dgops
