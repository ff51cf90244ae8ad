#include "../registry.hpp"

namespace dgops
{
  static OpRegistrar r0(op_key(3, 7, 4, 1, 0, 1, 0, 2, 0), &launch_op<3, 7, 4, float, 0, true, false, 2, 0>);
  static OpRegistrar r1(op_key(3, 7, 3, 0, 0, 1, 0, 2, 0), &launch_op<3, 7, 3, double, 0, true, false, 2, 0>);
  static OpRegistrar r2(op_key(3, 7, 2, 1, 0, 1, 0, 2, 0), &launch_op<3, 7, 2, float, 0, true, false, 2, 0>);
  static OpRegistrar r3(op_key(3, 8, 1, 1, 0, 1, 0, 1, 0), &launch_op<3, 8, 1, float, 0, true, false, 1, 0>);
  static OpRegistrar r4(op_key(3, 6, 2, 0, 0, 1, 0, 2, 0), &launch_op<3, 6, 2, double, 0, true, false, 2, 0>);
  static OpRegistrar r5(op_key(3, 5, 4, 0, 2, 1, 0, 2, 0), &launch_op<3, 5, 4, double, 2, true, false, 2, 0>);
  static MassRegistrar m6(mass_key(3, 7, 0), &launch_mass<3, 7, double>);
  static FdmRegistrar f7(fdm_key(3, 5, 3, 1), &launch_fdm<3, 5, 3, float, float, float>);
  static OpRegistrar r8(op_key(3, 6, 1, 0, 2, 1, 0, 1, 0), &launch_op<3, 6, 1, double, 2, true, false, 1, 0>);
  static CellGmresRegistrar q9(cell_gmres_key(3, 6, 3, 1), &launch_cell_gmres<3, 6, 3, float>);
  static FdmRegistrar f10(fdm_key(3, 5, 2, 2), &launch_fdm<3, 5, 2, float, double, double>);
  static OpRegistrar r11(op_key(2, 8, 1, 0, 2, 0, 1, 0, 1), &launch_op<2, 8, 1, double, 2, false, true, 0, 1>);
  static OpRegistrar r12(op_key(2, 6, 4, 1, 0, 1, 0, 2, 0), &launch_op<2, 6, 4, float, 0, true, false, 2, 0>);
  static OpRegistrar r13(op_key(2, 7, 1, 0, 1, 0, 1, 0, 0), &launch_op<2, 7, 1, double, 1, false, true, 0, 0>);
  static FdmRegistrar f14(fdm_key(2, 6, 3, 1), &launch_fdm<2, 6, 3, float, float, float>);
  static CellGmresRegistrar q15(cell_gmres_key(3, 5, 2, 0), &launch_cell_gmres<3, 5, 2, double>);
  static OpRegistrar r16(op_key(3, 4, 2, 1, 0, 1, 0, 2, 0), &launch_op<3, 4, 2, float, 0, true, false, 2, 0>);
  static OpRegistrar r17(op_key(2, 5, 4, 1, 0, 1, 0, 1, 0), &launch_op<2, 5, 4, float, 0, true, false, 1, 0>);
  static OpRegistrar r18(op_key(2, 6, 1, 1, 1, 0, 1, 0, 1), &launch_op<2, 6, 1, float, 1, false, true, 0, 1>);
  static OpRegistrar r19(op_key(2, 7, 1, 1, 2, 0, 0, 0, 0), &launch_op<2, 7, 1, float, 2, false, false, 0, 0>);
  static OpRegistrar r20(op_key(3, 3, 4, 0, 0, 1, 0, 1, 0), &launch_op<3, 3, 4, double, 0, true, false, 1, 0>);
  static OpRegistrar r21(op_key(2, 4, 4, 1, 2, 1, 0, 1, 0), &launch_op<2, 4, 4, float, 2, true, false, 1, 0>);
  static CellGmresRegistrar q22(cell_gmres_key(3, 4, 1, 1), &launch_cell_gmres<3, 4, 1, float>);
  static FdmRegistrar f23(fdm_key(2, 5, 2, 0), &launch_fdm<2, 5, 2, double, double, double>);
  static OpRegistrar r24(op_key(2, 6, 1, 0, 0, 1, 0, 2, 0), &launch_op<2, 6, 1, double, 0, true, false, 2, 0>);
  static OpRegistrar r25(op_key(2, 4, 3, 0, 2, 1, 0, 1, 0), &launch_op<2, 4, 3, double, 2, true, false, 1, 0>);
  static OpRegistrar r26(op_key(3, 3, 1, 1, 1, 0, 1, 0, 0), &launch_op<3, 3, 1, float, 1, false, true, 0, 0>);
  static OpRegistrar r27(op_key(2, 4, 1, 1, 1, 0, 1, 0, 1), &launch_op<2, 4, 1, float, 1, false, true, 0, 1>);
  static FdmRegistrar f28(fdm_key(2, 5, 1, 2), &launch_fdm<2, 5, 1, float, double, double>);
  static OpRegistrar r29(op_key(3, 3, 1, 0, 0, 0, 0, 0, 0), &launch_op<3, 3, 1, double, 0, false, false, 0, 0>);
  static FdmRegistrar f30(fdm_key(2, 3, 3, 2), &launch_fdm<2, 3, 3, float, double, double>);
  static OpRegistrar r31(op_key(2, 4, 1, 0, 2, 0, 0, 0, 1), &launch_op<2, 4, 1, double, 2, false, false, 0, 1>);
  static FdmRegistrar f32(fdm_key(3, 2, 4, 2), &launch_fdm<3, 2, 4, float, double, double>);
  static OpRegistrar r33(op_key(2, 3, 1, 0, 2, 0, 1, 0, 0), &launch_op<2, 3, 1, double, 2, false, true, 0, 0>);
  static OpRegistrar r34(op_key(2, 2, 4, 0, 2, 1, 0, 1, 0), &launch_op<2, 2, 4, double, 2, true, false, 1, 0>);
  static OpRegistrar r35(op_key(3, 2, 1, 1, 2, 0, 1, 0, 0), &launch_op<3, 2, 1, float, 2, false, true, 0, 0>);
  static FdmRegistrar f36(fdm_key(2, 3, 1, 1), &launch_fdm<2, 3, 1, float, float, float>);
  static OpRegistrar r37(op_key(2, 2, 3, 1, 0, 1, 0, 2, 0), &launch_op<2, 2, 3, float, 0, true, false, 2, 0>);
  static OpRegistrar r38(op_key(3, 2, 1, 0, 2, 1, 0, 2, 0), &launch_op<3, 2, 1, double, 2, true, false, 2, 0>);
  static CollocRegistrar c39(colloc_key(3, 2, 0), &launch_to_colloc<3, 2, double>);
  static FdmRegistrar f40(fdm_key(2, 2, 1, 0), &launch_fdm<2, 2, 1, double, double, double>);
} // namespace dgops
