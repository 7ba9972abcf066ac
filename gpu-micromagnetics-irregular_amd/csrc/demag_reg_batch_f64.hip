// demag_reg_batch_f64.hip -- the batched forms of the register-resident
// demag FFT passes (several instances of one plan per launch), double
// pipeline, in a translation unit of their own so that they compile in
// parallel with the single-instance kernels (demag_reg_f64.hip).
#include "demag_reg_kernels.h"

namespace demag_reg {
template int launch_row_fwd_batch<double>(int, int, int, int, const RegBatch<double>&, long, int,
                                          int, const C2<double>*, const C2<double>*, hipStream_t);
template int launch_col_batch<double>(int, int, int, int, const RegBatch<double>&, int,
                                      const C2<double>*, int, hipStream_t);
template int launch_row_inv_batch<double>(int, int, int, int, const RegBatch<double>&, int,
                                          const C2<double>*, const C2<double>*, long, int,
                                          hipStream_t);
}  // namespace demag_reg
