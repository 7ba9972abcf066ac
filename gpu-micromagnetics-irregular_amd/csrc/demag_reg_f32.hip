// demag_reg_f32.hip -- register-resident demag FFT passes, float pipeline
// (the kernels and launchers are in demag_reg_kernels.h).
#include "demag_reg_kernels.h"

namespace demag_reg {
template bool has_row_plan<float>(int, int);
template bool has_col_plan<float>(int, int);
template int launch_row_fwd<float>(int, int, int, const double*, const double*, long, int, int,
                                  const C2<float>*, const C2<float>*, C2<float>*, hipStream_t);
template int launch_col<float>(int, int, int, C2<float>*, int, const C2<float>*, const float*, int,
                              hipStream_t, hipEvent_t,
                              hipEvent_t);
template int launch_row_inv<float>(int, int, int, const C2<float>*, int, const C2<float>*,
                                  const C2<float>*, double*, long, int, hipStream_t);
}  // namespace demag_reg
