#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 4, 0), &launch_fdm<3, 8, 4, double, double, double>);
  static OpRegistrar r1(op_key(3, 8, 1, 0, 2, 0, 1, 0, 0), &launch_op<3, 8, 1, double, 2, false, true, 0, 0>);
  static OpRegistrar r2(op_key(3, 6, 4, 0, 2, 1, 0, 1, 0), &launch_op<3, 6, 4, double, 2, true, false, 1, 0>);
  static OpRegistrar r3(op_key(3, 7, 1, 1, 1, 0, 1, 0, 0), &launch_op<3, 7, 1, float, 1, false, true, 0, 0>);
  static MassRegistrar m4(mass_key(3, 8, 1), &launch_mass<3, 8, float>);
  static OpRegistrar r5(op_key(3, 6, 1, 0, 2, 0, 1, 0, 1), &launch_op<3, 6, 1, double, 2, false, true, 0, 1>);
  static OpRegistrar r6(op_key(3, 7, 1, 0, 2, 0, 0, 0, 0), &launch_op<3, 7, 1, double, 2, false, false, 0, 0>);
  static CellGmresRegistrar q7(cell_gmres_key(3, 7, 1, 1), &launch_cell_gmres<3, 7, 1, float>);
  static OpRegistrar r8(op_key(2, 7, 4, 1, 2, 1, 0, 1, 0), &launch_op<2, 7, 4, float, 2, true, false, 1, 0>);
  static FdmRegistrar f9(fdm_key(3, 6, 1, 2), &launch_fdm<3, 6, 1, float, double, double>);
  static OpRegistrar r10(op_key(3, 5, 2, 1, 0, 1, 0, 2, 0), &launch_op<3, 5, 2, float, 0, true, false, 2, 0>);
  static OpRegistrar r11(op_key(2, 7, 3, 1, 2, 1, 0, 2, 0), &launch_op<2, 7, 3, float, 2, true, false, 2, 0>);
  static FdmRegistrar f12(fdm_key(2, 8, 2, 0), &launch_fdm<2, 8, 2, double, double, double>);
  static OpRegistrar r13(op_key(3, 4, 3, 1, 2, 1, 0, 1, 0), &launch_op<3, 4, 3, float, 2, true, false, 1, 0>);
  static OpRegistrar r14(op_key(2, 7, 1, 1, 1, 1, 1, 1, 0), &launch_op<2, 7, 1, float, 1, true, true, 1, 0>);
  static OpRegistrar r15(op_key(3, 5, 1, 0, 0, 1, 0, 1, 0), &launch_op<3, 5, 1, double, 0, true, false, 1, 0>);
  static OpRegistrar r16(op_key(2, 8, 1, 0, 0, 0, 0, 0, 0), &launch_op<2, 8, 1, double, 0, false, false, 0, 0>);
  static OpRegistrar r17(op_key(3, 4, 1, 1, 1, 0, 1, 0, 0), &launch_op<3, 4, 1, float, 1, false, true, 0, 0>);
  static FdmRegistrar f18(fdm_key(2, 5, 4, 2), &launch_fdm<2, 5, 4, float, double, double>);
  static OpRegistrar r19(op_key(2, 5, 3, 0, 0, 1, 0, 1, 0), &launch_op<2, 5, 3, double, 0, true, false, 1, 0>);
  static OpRegistrar r20(op_key(2, 7, 1, 1, 2, 1, 0, 2, 0), &launch_op<2, 7, 1, float, 2, true, false, 2, 0>);
  static FdmRegistrar f21(fdm_key(3, 3, 4, 0), &launch_fdm<3, 3, 4, double, double, double>);
  static OpRegistrar r22(op_key(3, 4, 1, 1, 2, 1, 0, 2, 0), &launch_op<3, 4, 1, float, 2, true, false, 2, 0>);
  static OpRegistrar r23(op_key(2, 5, 2, 0, 0, 1, 0, 2, 0), &launch_op<2, 5, 2, double, 0, true, false, 2, 0>);
  static OpRegistrar r24(op_key(3, 3, 3, 0, 2, 1, 0, 2, 0), &launch_op<3, 3, 3, double, 2, true, false, 2, 0>);
  static BurgersRegistrar g25(burgers_key(2, 4, 0, 1), &launch_burgers<2, 4, 6, false, 1>);
  static FdmRegistrar f26(fdm_key(2, 4, 3, 1), &launch_fdm<2, 4, 3, float, float, float>);
  static FdmRegistrar f27(fdm_key(3, 3, 2, 2), &launch_fdm<3, 3, 2, float, double, double>);
  static OpRegistrar r28(op_key(2, 5, 1, 0, 0, 0, 0, 0, 0), &launch_op<2, 5, 1, double, 0, false, false, 0, 0>);
  static CollocRegistrar c29(colloc_key(2, 5, 1), &launch_to_colloc<2, 5, float>);
  static OpRegistrar r30(op_key(3, 3, 1, 0, 2, 1, 0, 1, 0), &launch_op<3, 3, 1, double, 2, true, false, 1, 0>);
  static BatchRegistrar s31(batch_key(3, 3), &launch_batch<3, 3>);
  static OpRegistrar r32(op_key(3, 2, 4, 0, 2, 1, 0, 1, 0), &launch_op<3, 2, 4, double, 2, true, false, 1, 0>);
  static CollocRegistrar c33(colloc_key(2, 4, 1), &launch_to_colloc<2, 4, float>);
  static OpRegistrar r34(op_key(2, 3, 1, 1, 2, 1, 1, 1, 0), &launch_op<2, 3, 1, float, 2, true, true, 1, 0>);
  static OpRegistrar r35(op_key(2, 2, 4, 0, 2, 1, 0, 2, 0), &launch_op<2, 2, 4, double, 2, true, false, 2, 0>);
  static FdmRegistrar f36(fdm_key(3, 2, 2, 2), &launch_fdm<3, 2, 2, float, double, double>);
  static CollocRegistrar c37(colloc_key(2, 3, 0), &launch_to_colloc<2, 3, double>);
  static OpRegistrar r38(op_key(3, 2, 1, 0, 2, 0, 0, 0, 1), &launch_op<3, 2, 1, double, 2, false, false, 0, 1>);
  static OpRegistrar r39(op_key(2, 2, 1, 0, 2, 1, 1, 1, 0), &launch_op<2, 2, 1, double, 2, true, true, 1, 0>);
  static CellGmresRegistrar q40(cell_gmres_key(3, 2, 3, 1), &launch_cell_gmres<3, 2, 3, float>);
  static CellGmresRegistrar q41(cell_gmres_key(2, 2, 3, 0), &launch_cell_gmres<2, 2, 3, double>);
} // namespace dgops
