// demag_reg_f64.hip -- register-resident demag FFT passes, double pipeline
// (the kernels and launchers are in demag_reg_kernels.h).
#include "demag_reg_kernels.h"

namespace demag_reg {
template bool has_row_plan<double>(int, int);
template bool has_col_plan<double>(int, int);
template int launch_row_fwd<double>(int, int, int, const double*, const double*, long, int, int,
                                  const C2<double>*, const C2<double>*, C2<double>*, hipStream_t);
template int launch_col<double>(int, int, int, C2<double>*, int, const C2<double>*, const double*, int,
                              hipStream_t, hipEvent_t,
                              hipEvent_t);
template int launch_row_inv<double>(int, int, int, const C2<double>*, int, const C2<double>*,
                                  const C2<double>*, double*, long, int, hipStream_t);
}  // namespace demag_reg
