#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 8, 3, 2), &launch_fdm<3, 8, 3, float, double, double>);
  static OpRegistrar r1(op_key(3, 8, 1, 1, 1, 1, 1, 1, 0), &launch_op<3, 8, 1, float, 1, true, true, 1, 0>);
  static OpRegistrar r2(op_key(3, 6, 4, 1, 0, 1, 0, 2, 0), &launch_op<3, 6, 4, float, 0, true, false, 2, 0>);
  static OpRegistrar r3(op_key(3, 7, 1, 1, 2, 1, 1, 1, 0), &launch_op<3, 7, 1, float, 2, true, true, 1, 0>);
  static OpRegistrar r4(op_key(3, 6, 3, 0, 0, 1, 0, 2, 0), &launch_op<3, 6, 3, double, 0, true, false, 2, 0>);
  static OpRegistrar r5(op_key(3, 6, 1, 1, 2, 0, 1, 0, 0), &launch_op<3, 6, 1, float, 2, false, true, 0, 0>);
  static OpRegistrar r6(op_key(3, 7, 1, 1, 0, 1, 0, 1, 0), &launch_op<3, 7, 1, float, 0, true, false, 1, 0>);
  static FdmRegistrar f7(fdm_key(2, 8, 4, 1), &launch_fdm<2, 8, 4, float, float, float>);
  static FdmRegistrar f8(fdm_key(2, 7, 4, 0), &launch_fdm<2, 7, 4, double, double, double>);
  static BatchRegistrar s9(batch_key(3, 6), &launch_batch<3, 6>);
  static OpRegistrar r10(op_key(3, 5, 1, 0, 2, 0, 1, 0, 0), &launch_op<3, 5, 1, double, 2, false, true, 0, 0>);
  static OpRegistrar r11(op_key(3, 4, 4, 0, 2, 1, 0, 1, 0), &launch_op<3, 4, 4, double, 2, true, false, 1, 0>);
  static FdmRegistrar f12(fdm_key(3, 4, 4, 2), &launch_fdm<3, 4, 4, float, double, double>);
  static FdmRegistrar f13(fdm_key(3, 4, 3, 0), &launch_fdm<3, 4, 3, double, double, double>);
  static FdmRegistrar f14(fdm_key(2, 7, 2, 1), &launch_fdm<2, 7, 2, float, float, float>);
  static OpRegistrar r15(op_key(3, 5, 1, 0, 2, 1, 0, 2, 0), &launch_op<3, 5, 1, double, 2, true, false, 2, 0>);
  static OpRegistrar r16(op_key(2, 8, 1, 1, 2, 0, 0, 0, 0), &launch_op<2, 8, 1, float, 2, false, false, 0, 0>);
  static OpRegistrar r17(op_key(3, 4, 1, 1, 2, 1, 1, 1, 0), &launch_op<3, 4, 1, float, 2, true, true, 1, 0>);
  static OpRegistrar r18(op_key(2, 6, 2, 0, 0, 1, 0, 2, 0), &launch_op<2, 6, 2, double, 0, true, false, 2, 0>);
  static OpRegistrar r19(op_key(2, 5, 3, 0, 2, 1, 0, 2, 0), &launch_op<2, 5, 3, double, 2, true, false, 2, 0>);
  static MassRegistrar m20(mass_key(2, 7, 1), &launch_mass<2, 7, float>);
  static OpRegistrar r21(op_key(3, 4, 1, 0, 2, 0, 0, 0, 0), &launch_op<3, 4, 1, double, 2, false, false, 0, 0>);
  static FdmRegistrar f22(fdm_key(3, 4, 1, 1), &launch_fdm<3, 4, 1, float, float, float>);
  static OpRegistrar r23(op_key(2, 5, 1, 0, 1, 0, 1, 0, 1), &launch_op<2, 5, 1, double, 1, false, true, 0, 1>);
  static FdmRegistrar f24(fdm_key(3, 3, 3, 2), &launch_fdm<3, 3, 3, float, double, double>);
  static CollocRegistrar c25(colloc_key(2, 6, 0), &launch_to_colloc<2, 6, double>);
  static OpRegistrar r26(op_key(3, 3, 2, 1, 2, 1, 0, 1, 0), &launch_op<3, 3, 2, float, 2, true, false, 1, 0>);
  static OpRegistrar r27(op_key(2, 4, 2, 0, 0, 1, 0, 2, 0), &launch_op<2, 4, 2, double, 0, true, false, 2, 0>);
  static OpRegistrar r28(op_key(2, 5, 1, 1, 2, 0, 0, 0, 0), &launch_op<2, 5, 1, float, 2, false, false, 0, 0>);
  static OpRegistrar r29(op_key(2, 3, 4, 0, 0, 1, 0, 1, 0), &launch_op<2, 3, 4, double, 0, true, false, 1, 0>);
  static OpRegistrar r30(op_key(2, 3, 3, 1, 2, 1, 0, 1, 0), &launch_op<2, 3, 3, float, 2, true, false, 1, 0>);
  static CellGmresRegistrar q31(cell_gmres_key(3, 3, 2, 0), &launch_cell_gmres<3, 3, 2, double>);
  static OpRegistrar r32(op_key(2, 4, 1, 1, 0, 1, 0, 2, 0), &launch_op<2, 4, 1, float, 0, true, false, 2, 0>);
  static OpRegistrar r33(op_key(2, 3, 2, 0, 0, 1, 0, 1, 0), &launch_op<2, 3, 2, double, 0, true, false, 1, 0>);
  static OpRegistrar r34(op_key(3, 2, 3, 0, 2, 1, 0, 1, 0), &launch_op<3, 2, 3, double, 2, true, false, 1, 0>);
  static OpRegistrar r35(op_key(3, 2, 2, 1, 0, 1, 0, 2, 0), &launch_op<3, 2, 2, float, 0, true, false, 2, 0>);
  static OpRegistrar r36(op_key(2, 3, 1, 1, 0, 0, 0, 0, 0), &launch_op<2, 3, 1, float, 0, false, false, 0, 0>);
  static CollocRegistrar c37(colloc_key(2, 3, 1), &launch_to_colloc<2, 3, float>);
  static OpRegistrar r38(op_key(3, 2, 1, 0, 2, 1, 0, 1, 0), &launch_op<3, 2, 1, double, 2, true, false, 1, 0>);
  static OpRegistrar r39(op_key(2, 2, 1, 1, 2, 0, 1, 0, 1), &launch_op<2, 2, 1, float, 2, false, true, 0, 1>);
  static OpRegistrar r40(op_key(2, 2, 1, 0, 2, 0, 0, 0, 1), &launch_op<2, 2, 1, double, 2, false, false, 0, 1>);
  static CellGmresRegistrar q41(cell_gmres_key(2, 2, 3, 1), &launch_cell_gmres<2, 2, 3, float>);
} // namespace dgops
