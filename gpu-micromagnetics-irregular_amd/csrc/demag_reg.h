// demag_reg.h -- launchers of the register-resident demag FFT passes
// (csrc/demag_reg_kernels.h), one instantiation per pipeline precision
// (demag_reg_f64.hip, demag_reg_f32.hip).  Internal to the library.
//
// Each launcher returns 1 when it launched, 0 when no register plan covers
// the length (the caller then runs the general LDS Stockham kernels) and
// -1 when the launch failed.  `half` selects the zero-padded form: only the
// first half of a transform's inputs is non-zero and only the first half of
// its outputs is kept.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_regs.h"

namespace demag_reg {

template <class T>
using C2 = rfft::Cplx<T>;

// A batch runs one pass for up to kBatchMax instances that share a plan
// (same lengths, precision and variant) in one launch: blockIdx.y selects
// the instance, whose buffers come from these arrays (kernel arguments,
// copied at launch, so the caller may reuse them at once).  Each
// instance's workgroups do exactly the work of its own launch -- same
// plan, same lanes, same operations -- so every instance gets the bits of
// its single launch.  The twiddle and post-processing tables depend only
// on the plan and are taken once for all.
constexpr int kBatchMax = 32;
template <class T>
struct RegBatch {
  const double* in[kBatchMax];   // row forward: SoA state
  const double* w[kBatchMax];    // row forward: window, or null
  C2<T>* spec[kBatchMax];        // intermediate spectrum
  const T* K[kBatchMax];         // column: tensor planes
  double* out[kBatchMax];        // row inverse: SoA field (overwrite)
};

// true when register plan `variant` exists for a row (complex length
// L = Px/2) or a column (length L = Py) transform.  Rows have variant 0;
// columns have variant 0 (8 values per component and lane) for every
// length, variant 1 (16 values per lane) for lengths 1024 to 8192, and
// variants 2, 3, 4 (variant 0's transforms with 2, 4 or 8 adjacent columns
// per workgroup and the spectrum moved in runs of that many values) for
// lengths 512 to 2048 where the block fits one workgroup, and variants 5, 6
// (variant 1's transforms, one per wave, in workgroups of 2 or 4 columns'
// x/y or 4 or 8 columns' z) for length 1024.
template <class T> bool has_row_plan(int L, int variant);
template <class T> bool has_col_plan(int L, int variant);

// row pass, forward: per (row, component) gather of the SoA state (times the
// window w when given), real FFT of length 2L, L + 1 bins to spec
template <class T>
int launch_row_fwd(int L, int variant, int half, const double* in, const double* w, long comp_stride,
                   int nx_in, int rows, const C2<T>* tw, const C2<T>* post, C2<T>* spec,
                   hipStream_t s);

// column pass: forward column FFTs of the three components, 3x3 tensor
// product (planes xx, yy, zz, xy-real, xy-imag stored [kx][ky]), inverse
// column FFTs; spec is [comp][row][kx] with row count L (L/2 when half).
// e0/e1 (may be null): events the dispatch itself records at the start and
// the end of the kernel (timing without separate marker packets)
template <class T>
int launch_col(int L, int variant, int half, C2<T>* spec, int Hx, const C2<T>* tw, const T* K, int has_xy_im,
               hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);

// row pass, inverse: inverse real FFT per (row, component), field written
// into the SoA output (overwrite)
template <class T>
int launch_row_inv(int L, int variant, int half, const C2<T>* spec, int rows, const C2<T>* tw,
                   const C2<T>* post, double* out, long comp_stride, int nx_out, hipStream_t s);

// batched forms of the three passes (FP64 pipeline): n <= kBatchMax
// instances of one plan, buffers from b; same return convention
template <class T>
int launch_row_fwd_batch(int L, int variant, int half, int n, const RegBatch<T>& b,
                         long comp_stride, int nx_in, int rows, const C2<T>* tw,
                         const C2<T>* post, hipStream_t s);
template <class T>
int launch_col_batch(int L, int variant, int half, int n, const RegBatch<T>& b, int Hx,
                     const C2<T>* tw, int has_xy_im, hipStream_t s);
template <class T>
int launch_row_inv_batch(int L, int variant, int half, int n, const RegBatch<T>& b, int rows,
                         const C2<T>* tw, const C2<T>* post, long comp_stride, int nx_out,
                         hipStream_t s);

extern template int launch_row_fwd_batch<double>(int, int, int, int, const RegBatch<double>&,
                                                 long, int, int, const C2<double>*,
                                                 const C2<double>*, hipStream_t);
extern template int launch_col_batch<double>(int, int, int, int, const RegBatch<double>&, int,
                                             const C2<double>*, int, hipStream_t);
extern template int launch_row_inv_batch<double>(int, int, int, int, const RegBatch<double>&, int,
                                                 const C2<double>*, const C2<double>*, long, int,
                                                 hipStream_t);

extern template bool has_row_plan<double>(int, int);
extern template bool has_col_plan<double>(int, int);
extern template bool has_row_plan<float>(int, int);
extern template bool has_col_plan<float>(int, int);
extern template int launch_row_fwd<double>(int, int, int, const double*, const double*, long, int, int,
                                           const C2<double>*, const C2<double>*, C2<double>*,
                                           hipStream_t);
extern template int launch_row_fwd<float>(int, int, int, const double*, const double*, long, int, int,
                                          const C2<float>*, const C2<float>*, C2<float>*,
                                          hipStream_t);
extern template int launch_col<double>(int, int, int, C2<double>*, int, const C2<double>*,
                                       const double*, int, hipStream_t, hipEvent_t, hipEvent_t);
extern template int launch_col<float>(int, int, int, C2<float>*, int, const C2<float>*, const float*,
                                      int, hipStream_t, hipEvent_t, hipEvent_t);
extern template int launch_row_inv<double>(int, int, int, const C2<double>*, int, const C2<double>*,
                                           const C2<double>*, double*, long, int, hipStream_t);
extern template int launch_row_inv<float>(int, int, int, const C2<float>*, int, const C2<float>*,
                                          const C2<float>*, double*, long, int, hipStream_t);

}  // namespace demag_reg
