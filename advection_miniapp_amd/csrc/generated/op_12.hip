#include "../registry.hpp"

namespace dgops
{
  static FdmRegistrar f0(fdm_key(3, 7, 4, 0), &launch_fdm<3, 7, 4, double, double, double>);
  static OpRegistrar r1(op_key(3, 7, 1, 0, 1, 0, 1, 0, 0), &launch_op<3, 7, 1, double, true, false, true, 0, 0>);
  static FdmRegistrar f2(fdm_key(3, 6, 3, 1), &launch_fdm<3, 6, 3, float, float, float>);
  static OpRegistrar r3(op_key(3, 7, 1, 1, 1, 0, 0, 0, 0), &launch_op<3, 7, 1, float, true, false, false, 0, 0>);
  static FdmRegistrar f4(fdm_key(2, 7, 4, 1), &launch_fdm<2, 7, 4, float, float, float>);
  static OpRegistrar r5(op_key(3, 5, 1, 0, 1, 1, 1, 1, 0), &launch_op<3, 5, 1, double, true, true, true, 1, 0>);
  static FdmRegistrar f6(fdm_key(3, 4, 4, 1), &launch_fdm<3, 4, 4, float, float, float>);
  static FdmRegistrar f7(fdm_key(2, 7, 2, 0), &launch_fdm<2, 7, 2, double, double, double>);
  static OpRegistrar r8(op_key(2, 8, 1, 1, 1, 0, 0, 0, 0), &launch_op<2, 8, 1, float, true, false, false, 0, 0>);
  static OpRegistrar r9(op_key(2, 6, 2, 0, 1, 1, 0, 1, 0), &launch_op<2, 6, 2, double, true, true, false, 1, 0>);
  static OpRegistrar r10(op_key(2, 7, 1, 1, 1, 1, 0, 1, 0), &launch_op<2, 7, 1, float, true, true, false, 1, 0>);
  static FdmRegistrar f11(fdm_key(3, 4, 1, 1), &launch_fdm<3, 4, 1, float, float, float>);
  static OpRegistrar r12(op_key(2, 6, 1, 1, 0, 1, 0, 1, 0), &launch_op<2, 6, 1, float, false, true, false, 1, 0>);
  static FdmRegistrar f13(fdm_key(3, 3, 2, 1), &launch_fdm<3, 3, 2, float, float, float>);
  static FdmRegistrar f14(fdm_key(2, 5, 1, 2), &launch_fdm<2, 5, 1, float, double, double>);
  static BurgersRegistrar g15(burgers_key(3, 3, 0, 1), &launch_burgers<3, 3, 3, false, 1>);
  static OpRegistrar r16(op_key(2, 3, 1, 0, 1, 1, 1, 1, 0), &launch_op<2, 3, 1, double, true, true, true, 1, 0>);
  static FdmRegistrar f17(fdm_key(2, 2, 4, 2), &launch_fdm<2, 2, 4, float, double, double>);
  static FdmRegistrar f18(fdm_key(2, 2, 3, 2), &launch_fdm<2, 2, 3, float, double, double>);
  static OpRegistrar r19(op_key(2, 2, 1, 0, 0, 0, 0, 0, 0), &launch_op<2, 2, 1, double, false, false, false, 0, 0>);
} // namespace dgops
