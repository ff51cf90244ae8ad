#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 8, 1, 0, 1, 1, 1, 1, 0), &launch_op<3, 8, 1, double, 1, true, true, 1, 0>);
  static OpRegistrar r1(op_key(3, 6, 4, 0, 0, 1, 0, 1, 0), &launch_op<3, 6, 4, double, 0, true, false, 1, 0>);
  static OpRegistrar r2(op_key(3, 7, 1, 0, 2, 1, 1, 1, 0), &launch_op<3, 7, 1, double, 2, true, true, 1, 0>);
  static MassRegistrar m3(mass_key(3, 8, 0), &launch_mass<3, 8, double>);
  static OpRegistrar r4(op_key(3, 6, 1, 0, 2, 0, 1, 0, 0), &launch_op<3, 6, 1, double, 2, false, true, 0, 0>);
  static OpRegistrar r5(op_key(3, 7, 1, 0, 0, 0, 0, 0, 1), &launch_op<3, 7, 1, double, 0, false, false, 0, 1>);
  static CollocRegistrar c6(colloc_key(3, 7, 1), &launch_to_colloc<3, 7, float>);
  static OpRegistrar r7(op_key(2, 7, 4, 1, 0, 1, 0, 1, 0), &launch_op<2, 7, 4, float, 0, true, false, 1, 0>);
  static FdmRegistrar f8(fdm_key(3, 6, 1, 1), &launch_fdm<3, 6, 1, float, float, float>);
  static OpRegistrar r9(op_key(3, 5, 2, 0, 2, 1, 0, 2, 0), &launch_op<3, 5, 2, double, 2, true, false, 2, 0>);
  static OpRegistrar r10(op_key(2, 7, 3, 1, 0, 1, 0, 2, 0), &launch_op<2, 7, 3, float, 0, true, false, 2, 0>);
  static OpRegistrar r11(op_key(2, 8, 1, 1, 2, 1, 1, 1, 0), &launch_op<2, 8, 1, float, 2, true, true, 1, 0>);
  static OpRegistrar r12(op_key(3, 4, 3, 1, 0, 1, 0, 1, 0), &launch_op<3, 4, 3, float, 0, true, false, 1, 0>);
  static OpRegistrar r13(op_key(2, 7, 1, 1, 1, 0, 1, 0, 1), &launch_op<2, 7, 1, float, 1, false, true, 0, 1>);
  static OpRegistrar r14(op_key(3, 5, 1, 1, 2, 0, 0, 0, 0), &launch_op<3, 5, 1, float, 2, false, false, 0, 0>);
  static CellGmresRegistrar q15(cell_gmres_key(3, 5, 4, 1), &launch_cell_gmres<3, 5, 4, float>);
  static OpRegistrar r16(op_key(3, 4, 1, 0, 2, 1, 1, 1, 0), &launch_op<3, 4, 1, double, 2, true, true, 1, 0>);
  static FdmRegistrar f17(fdm_key(2, 5, 4, 1), &launch_fdm<2, 5, 4, float, float, float>);
  static FdmRegistrar f18(fdm_key(2, 6, 2, 2), &launch_fdm<2, 6, 2, float, double, double>);
  static OpRegistrar r19(op_key(2, 7, 1, 1, 0, 1, 0, 2, 0), &launch_op<2, 7, 1, float, 0, true, false, 2, 0>);
  static OpRegistrar r20(op_key(3, 3, 4, 1, 2, 1, 0, 2, 0), &launch_op<3, 3, 4, float, 2, true, false, 2, 0>);
  static OpRegistrar r21(op_key(3, 4, 1, 1, 0, 1, 0, 2, 0), &launch_op<3, 4, 1, float, 0, true, false, 2, 0>);
  static OpRegistrar r22(op_key(2, 5, 2, 1, 2, 1, 0, 1, 0), &launch_op<2, 5, 2, float, 2, true, false, 1, 0>);
  static OpRegistrar r23(op_key(3, 3, 3, 0, 0, 1, 0, 2, 0), &launch_op<3, 3, 3, double, 0, true, false, 2, 0>);
  static BurgersRegistrar g24(burgers_key(2, 4, 0, 0), &launch_burgers<2, 4, 6, false, 0>);
  static FdmRegistrar f25(fdm_key(2, 4, 3, 0), &launch_fdm<2, 4, 3, double, double, double>);
  static FdmRegistrar f26(fdm_key(3, 3, 2, 1), &launch_fdm<3, 3, 2, float, float, float>);
  static FdmRegistrar f27(fdm_key(2, 4, 2, 2), &launch_fdm<2, 4, 2, float, double, double>);
  static CellGmresRegistrar q28(cell_gmres_key(2, 5, 4, 0), &launch_cell_gmres<2, 5, 4, double>);
  static OpRegistrar r29(op_key(3, 3, 1, 1, 0, 1, 0, 1, 0), &launch_op<3, 3, 1, float, 0, true, false, 1, 0>);
  static BurgersRegistrar g30(burgers_key(3, 3, 1, 1), &launch_burgers<3, 3, 3, true, 1>);
  static OpRegistrar r31(op_key(3, 2, 4, 0, 0, 1, 0, 1, 0), &launch_op<3, 2, 4, double, 0, true, false, 1, 0>);
  static CellGmresRegistrar q32(cell_gmres_key(2, 4, 4, 0), &launch_cell_gmres<2, 4, 4, double>);
  static OpRegistrar r33(op_key(2, 3, 1, 1, 2, 0, 1, 0, 1), &launch_op<2, 3, 1, float, 2, false, true, 0, 1>);
  static OpRegistrar r34(op_key(2, 2, 4, 0, 0, 1, 0, 2, 0), &launch_op<2, 2, 4, double, 0, true, false, 2, 0>);
  static FdmRegistrar f35(fdm_key(3, 2, 2, 1), &launch_fdm<3, 2, 2, float, float, float>);
  static MassRegistrar m36(mass_key(2, 3, 1), &launch_mass<2, 3, float>);
  static OpRegistrar r37(op_key(3, 2, 1, 0, 2, 0, 0, 0, 0), &launch_op<3, 2, 1, double, 2, false, false, 0, 0>);
  static OpRegistrar r38(op_key(2, 2, 1, 0, 2, 0, 1, 0, 1), &launch_op<2, 2, 1, double, 2, false, true, 0, 1>);
  static CellGmresRegistrar q39(cell_gmres_key(3, 2, 2, 1), &launch_cell_gmres<3, 2, 2, float>);
  static CellGmresRegistrar q40(cell_gmres_key(2, 2, 2, 0), &launch_cell_gmres<2, 2, 2, double>);
} // namespace dgops
