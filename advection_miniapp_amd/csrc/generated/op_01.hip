#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 4, 1), &launch_fdm<3, 8, 4, float, float, float>);
  static OpRegistrar r1(op_key(3, 8, 1, 0, 2, 0, 1, 0, 1), &launch_op<3, 8, 1, double, 2, false, true, 0, 1>);
  static OpRegistrar r2(op_key(3, 6, 4, 1, 0, 1, 0, 1, 0), &launch_op<3, 6, 4, float, 0, true, false, 1, 0>);
  static OpRegistrar r3(op_key(3, 7, 1, 1, 1, 0, 1, 0, 1), &launch_op<3, 7, 1, float, 1, false, true, 0, 1>);
  static OpRegistrar r4(op_key(3, 6, 3, 0, 0, 1, 0, 1, 0), &launch_op<3, 6, 3, double, 0, true, false, 1, 0>);
  static OpRegistrar r5(op_key(3, 6, 1, 0, 2, 1, 1, 1, 0), &launch_op<3, 6, 1, double, 2, true, true, 1, 0>);
  static OpRegistrar r6(op_key(3, 7, 1, 0, 2, 0, 0, 0, 1), &launch_op<3, 7, 1, double, 2, false, false, 0, 1>);
  static CellGmresRegistrar q7(cell_gmres_key(3, 7, 2, 1), &launch_cell_gmres<3, 7, 2, float>);
  static OpRegistrar r8(op_key(2, 7, 4, 0, 0, 1, 0, 2, 0), &launch_op<2, 7, 4, double, 0, true, false, 2, 0>);
  static BurgersRegistrar g9(burgers_key(3, 4, 0, 0), &launch_burgers<3, 4, 6, false, 0>);
  static OpRegistrar r10(op_key(3, 5, 2, 1, 2, 1, 0, 2, 0), &launch_op<3, 5, 2, float, 2, true, false, 2, 0>);
  static FdmRegistrar f11(fdm_key(2, 7, 3, 0), &launch_fdm<2, 7, 3, double, double, double>);
  static FdmRegistrar f12(fdm_key(2, 8, 2, 1), &launch_fdm<2, 8, 2, float, float, float>);
  static OpRegistrar r13(op_key(3, 4, 3, 0, 0, 1, 0, 2, 0), &launch_op<3, 4, 3, double, 0, true, false, 2, 0>);
  static OpRegistrar r14(op_key(2, 7, 1, 1, 2, 0, 1, 0, 0), &launch_op<2, 7, 1, float, 2, false, true, 0, 0>);
  static OpRegistrar r15(op_key(3, 5, 1, 1, 0, 1, 0, 1, 0), &launch_op<3, 5, 1, float, 0, true, false, 1, 0>);
  static OpRegistrar r16(op_key(2, 8, 1, 0, 0, 0, 0, 0, 1), &launch_op<2, 8, 1, double, 0, false, false, 0, 1>);
  static OpRegistrar r17(op_key(3, 4, 1, 1, 1, 0, 1, 0, 1), &launch_op<3, 4, 1, float, 1, false, true, 0, 1>);
  static OpRegistrar r18(op_key(2, 6, 2, 0, 0, 1, 0, 1, 0), &launch_op<2, 6, 2, double, 0, true, false, 1, 0>);
  static OpRegistrar r19(op_key(2, 5, 3, 0, 2, 1, 0, 1, 0), &launch_op<2, 5, 3, double, 2, true, false, 1, 0>);
  static FdmRegistrar f20(fdm_key(2, 7, 1, 0), &launch_fdm<2, 7, 1, double, double, double>);
  static FdmRegistrar f21(fdm_key(3, 3, 4, 1), &launch_fdm<3, 3, 4, float, float, float>);
  static FdmRegistrar f22(fdm_key(2, 4, 4, 0), &launch_fdm<2, 4, 4, double, double, double>);
  static OpRegistrar r23(op_key(2, 5, 2, 0, 2, 1, 0, 2, 0), &launch_op<2, 5, 2, double, 2, true, false, 2, 0>);
  static OpRegistrar r24(op_key(3, 3, 3, 1, 0, 1, 0, 2, 0), &launch_op<3, 3, 3, float, 0, true, false, 2, 0>);
  static BurgersRegistrar g25(burgers_key(2, 4, 1, 0), &launch_burgers<2, 4, 6, true, 0>);
  static FdmRegistrar f26(fdm_key(2, 4, 3, 2), &launch_fdm<2, 4, 3, float, double, double>);
  static OpRegistrar r27(op_key(2, 4, 2, 0, 0, 1, 0, 1, 0), &launch_op<2, 4, 2, double, 0, true, false, 1, 0>);
  static OpRegistrar r28(op_key(2, 5, 1, 0, 0, 0, 0, 0, 1), &launch_op<2, 5, 1, double, 0, false, false, 0, 1>);
  static CellGmresRegistrar q29(cell_gmres_key(2, 5, 1, 1), &launch_cell_gmres<2, 5, 1, float>);
  static OpRegistrar r30(op_key(3, 3, 1, 1, 2, 1, 0, 1, 0), &launch_op<3, 3, 1, float, 2, true, false, 1, 0>);
  static MassRegistrar m31(mass_key(3, 3, 0), &launch_mass<3, 3, double>);
  static OpRegistrar r32(op_key(3, 2, 4, 1, 0, 1, 0, 1, 0), &launch_op<3, 2, 4, float, 0, true, false, 1, 0>);
  static CellGmresRegistrar q33(cell_gmres_key(2, 4, 1, 1), &launch_cell_gmres<2, 4, 1, float>);
  static FdmRegistrar f34(fdm_key(2, 3, 2, 0), &launch_fdm<2, 3, 2, double, double, double>);
  static OpRegistrar r35(op_key(2, 2, 4, 1, 0, 1, 0, 2, 0), &launch_op<2, 2, 4, float, 0, true, false, 2, 0>);
  static OpRegistrar r36(op_key(2, 3, 1, 0, 0, 0, 0, 0, 0), &launch_op<2, 3, 1, double, 0, false, false, 0, 0>);
  static CellGmresRegistrar q37(cell_gmres_key(2, 3, 1, 0), &launch_cell_gmres<2, 3, 1, double>);
  static OpRegistrar r38(op_key(3, 2, 1, 1, 0, 0, 0, 0, 0), &launch_op<3, 2, 1, float, 0, false, false, 0, 0>);
  static OpRegistrar r39(op_key(2, 2, 1, 1, 1, 0, 1, 0, 0), &launch_op<2, 2, 1, float, 1, false, true, 0, 0>);
  static CellGmresRegistrar q40(cell_gmres_key(3, 2, 4, 1), &launch_cell_gmres<3, 2, 4, float>);
  static CellGmresRegistrar q41(cell_gmres_key(2, 2, 4, 0), &launch_cell_gmres<2, 2, 4, double>);
} // namespace dgops
