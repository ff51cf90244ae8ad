#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 7, 4, 2), &launch_fdm<3, 7, 4, float, double, double>);
  static FdmRegistrar f1(fdm_key(3, 7, 3, 0), &launch_fdm<3, 7, 3, double, double, double>);
  static OpRegistrar r2(op_key(3, 7, 1, 0, 1, 1, 1, 1, 0), &launch_op<3, 7, 1, double, 1, true, true, 1, 0>);
  static FdmRegistrar f3(fdm_key(3, 8, 1, 1), &launch_fdm<3, 8, 1, float, float, float>);
  static OpRegistrar r4(op_key(3, 6, 1, 0, 1, 0, 1, 0, 0), &launch_op<3, 6, 1, double, 1, false, true, 0, 0>);
  static FdmRegistrar f5(fdm_key(3, 5, 4, 1), &launch_fdm<3, 5, 4, float, float, float>);
  static CellGmresRegistrar q6(cell_gmres_key(3, 7, 2, 0), &launch_cell_gmres<3, 7, 2, double>);
  static FdmRegistrar f7(fdm_key(2, 8, 3, 2), &launch_fdm<2, 8, 3, float, double, double>);
  static OpRegistrar r8(op_key(3, 6, 1, 1, 0, 1, 0, 2, 0), &launch_op<3, 6, 1, float, 0, true, false, 2, 0>);
  static OpRegistrar r9(op_key(3, 5, 2, 1, 0, 1, 0, 1, 0), &launch_op<3, 5, 2, float, 0, true, false, 1, 0>);
  static OpRegistrar r10(op_key(2, 7, 3, 1, 2, 1, 0, 1, 0), &launch_op<2, 7, 3, float, 2, true, false, 1, 0>);
  static OpRegistrar r11(op_key(2, 8, 1, 1, 1, 1, 1, 1, 0), &launch_op<2, 8, 1, float, 1, true, true, 1, 0>);
  static FdmRegistrar f12(fdm_key(2, 6, 4, 2), &launch_fdm<2, 6, 4, float, double, double>);
  static OpRegistrar r13(op_key(2, 7, 1, 0, 2, 0, 1, 0, 1), &launch_op<2, 7, 1, double, 2, false, true, 0, 1>);
  static OpRegistrar r14(op_key(3, 5, 1, 0, 2, 0, 0, 0, 0), &launch_op<3, 5, 1, double, 2, false, false, 0, 0>);
  static CellGmresRegistrar q15(cell_gmres_key(3, 5, 1, 1), &launch_cell_gmres<3, 5, 1, float>);
  static OpRegistrar r16(op_key(3, 4, 1, 0, 1, 1, 1, 1, 0), &launch_op<3, 4, 1, double, 1, true, true, 1, 0>);
  static OpRegistrar r17(op_key(2, 5, 4, 1, 0, 1, 0, 2, 0), &launch_op<2, 5, 4, float, 0, true, false, 2, 0>);
  static OpRegistrar r18(op_key(2, 6, 1, 1, 2, 1, 1, 1, 0), &launch_op<2, 6, 1, float, 2, true, true, 1, 0>);
  static OpRegistrar r19(op_key(2, 7, 1, 1, 2, 1, 0, 1, 0), &launch_op<2, 7, 1, float, 2, true, false, 1, 0>);
  static OpRegistrar r20(op_key(3, 3, 4, 0, 0, 1, 0, 2, 0), &launch_op<3, 3, 4, double, 0, true, false, 2, 0>);
  static OpRegistrar r21(op_key(2, 4, 4, 1, 2, 1, 0, 2, 0), &launch_op<2, 4, 4, float, 2, true, false, 2, 0>);
  static OpRegistrar r22(op_key(2, 5, 2, 0, 0, 1, 0, 1, 0), &launch_op<2, 5, 2, double, 0, true, false, 1, 0>);
  static OpRegistrar r23(op_key(3, 3, 3, 0, 2, 1, 0, 1, 0), &launch_op<3, 3, 3, double, 2, true, false, 1, 0>);
  static FdmRegistrar f24(fdm_key(2, 6, 1, 0), &launch_fdm<2, 6, 1, double, double, double>);
  static OpRegistrar r25(op_key(2, 4, 3, 0, 2, 1, 0, 2, 0), &launch_op<2, 4, 3, double, 2, true, false, 2, 0>);
  static OpRegistrar r26(op_key(3, 3, 1, 1, 2, 0, 1, 0, 1), &launch_op<3, 3, 1, float, 2, false, true, 0, 1>);
  static OpRegistrar r27(op_key(2, 4, 1, 1, 2, 1, 1, 1, 0), &launch_op<2, 4, 1, float, 2, true, true, 1, 0>);
  static CellGmresRegistrar q28(cell_gmres_key(2, 5, 1, 0), &launch_cell_gmres<2, 5, 1, double>);
  static OpRegistrar r29(op_key(3, 3, 1, 1, 0, 0, 0, 0, 0), &launch_op<3, 3, 1, float, 0, false, false, 0, 0>);
  static BurgersRegistrar g30(burgers_key(3, 3, 0, 0), &launch_burgers<3, 3, 3, false, 0>);
  static OpRegistrar r31(op_key(2, 4, 1, 1, 0, 1, 0, 1, 0), &launch_op<2, 4, 1, float, 0, true, false, 1, 0>);
  static CellGmresRegistrar q32(cell_gmres_key(2, 4, 1, 0), &launch_cell_gmres<2, 4, 1, double>);
  static OpRegistrar r33(op_key(2, 3, 1, 1, 1, 0, 1, 0, 1), &launch_op<2, 3, 1, float, 1, false, true, 0, 1>);
  static OpRegistrar r34(op_key(3, 2, 2, 0, 2, 1, 0, 1, 0), &launch_op<3, 2, 2, double, 2, true, false, 1, 0>);
  static FdmRegistrar f35(fdm_key(2, 2, 4, 1), &launch_fdm<2, 2, 4, float, float, float>);
  static BurgersRegistrar g36(burgers_key(2, 3, 1, 0), &launch_burgers<2, 3, 3, true, 0>);
  static FdmRegistrar f37(fdm_key(2, 2, 3, 2), &launch_fdm<2, 2, 3, float, double, double>);
  static OpRegistrar r38(op_key(2, 2, 1, 0, 1, 0, 1, 0, 1), &launch_op<2, 2, 1, double, 1, false, true, 0, 1>);
  static CellGmresRegistrar q39(cell_gmres_key(3, 2, 4, 0), &launch_cell_gmres<3, 2, 4, double>);
  static MassRegistrar m40(mass_key(2, 2, 1), &launch_mass<2, 2, float>);
} // namespace dgops
