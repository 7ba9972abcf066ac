/*
 * demag_fft.h -- FFT demagnetising field on gfx950, hand-written LDS radix
 * FFT (no vendor FFT library on the hot path).
 *
 * Replaces the reference module API src/irregular/2d_fft/2d_i4/demag_fft.h:29-42
 * (Demag_Init / Demag_Apply / Demag_GetSelfCoupling / Demag_Destroy) and the
 * polycrystal additions src/irregular/2d_fft/2d_i3/demag_fft.h
 * (Demag_ApplyWindowed, Demag_SetStream).
 *
 * Operator (reference semantics, 2d_i4/demag_fft.cu:64-91,143-161,325-363):
 *     h(r) = strength * sum_r' N(r - r') m(r')
 * with N the 81-point face-averaged Newell cell tensor (`calt`/`ctt`).
 *   Demag_Init       : periodic circular convolution on the nx x ny cell
 *                      grid; displacement wrapped into [-n/2, n - n/2).
 *   Demag_InitPadded : open-boundary (non-wrapped) convolution, evaluated
 *                      as a circular one on the zero-padded 2nx x 2ny grid.
 *                      No reference counterpart (the padded R2C of the
 *                      1536x512 square-hole workload).
 * y_dev and h_out_dev are SoA [mx|my|mz] device arrays of 3*nx*ny doubles;
 * h_out_dev is OVERWRITTEN (same contract as the reference).
 */
#ifndef LLG_DEMAG_FFT_H
#define LLG_DEMAG_FFT_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct DemagData DemagData;

DemagData* Demag_Init(int nx, int ny, double thick, double demag_strength);
DemagData* Demag_InitPadded(int nx, int ny, double thick, double demag_strength);

void Demag_Apply(DemagData* d, const double* y_dev, double* h_out_dev);
/* gather multiplies m by the per-cell weight w_dev[nx*ny] (2d_i3/demag_fft.cu:125-139) */
void Demag_ApplyWindowed(DemagData* d, const double* y_dev, const double* w_dev,
                         double* h_out_dev);
/* stream is a hipStream_t; NULL selects the library default stream */
void Demag_SetStream(DemagData* d, void* stream);
/* n fields in one set of launches on `stream`: h[i] = demag of y[i] (times
 * w[i] when w and w[i] are non-NULL), each exactly what Demag_Apply(d[i],
 * ...) writes.  Instances with the same FFT plan (grid, padding, FP64,
 * register kernel variants) share one launch per pass, up to 32 at a time;
 * the others run their own passes on `stream`.  No reference counterpart
 * (one grid per process there).  Returns 0, or -1 on a launch error. */
int  Demag_ApplyBatch(int n, DemagData* const* d, const double* const* y,
                      const double* const* w, double* const* h, void* stream);
/* N_aa(0) * strength; zeros when d == NULL (2d_i4/demag_fft.cu:365-377) */
void Demag_GetSelfCoupling(DemagData* d, double* nxx0, double* nyy0, double* nzz0);
void Demag_Destroy(DemagData* d);

/* FP32 spectral pipeline (FFT, tensor multiply and intermediates in FP32;
 * input gather and output field stay FP64).  fp32 = 0 restores FP64. */
int  Demag_SetPrecision(DemagData* d, int fp32);

/* LDS padding of the FFT passes: one spare complex slot per 2^shift values
 * (31 disables it) for the row and the column plans.  A tuning knob:
 * results do not depend on it. */
int  Demag_SetLdsPadding(DemagData* d, int row_shift, int col_shift);
/* Workgroup sizes of the row and column FFT kernels (multiples of 64 up to
 * 512; 0 keeps the current one).  Fails (-1, nothing changed for that
 * plan) when a pass would exceed its per-thread register budget.  A
 * tuning knob: results do not depend on it. */
int  Demag_SetWorkgroups(DemagData* d, int row_block, int col_block);
/* Compile-time FFT plans (on by default): kernels specialised for the
 * zero-padded 3072 x 1024 grid (rows of 1536 points, columns of 1024) and
 * the periodic 512 x 512 grid (rows of 256, columns of 512), at the
 * default workgroup size and padding of each plan, run whenever the plan
 * matches; other plans take the general kernels.  Same arithmetic, so
 * results do not depend on it.  Returns the number of the two transform
 * directions (row, column) of the current precision that take a
 * compile-time plan, or -1. */
int  Demag_SetSpecialized(DemagData* d, int on);
/* FFT kernel family per pass.  Variant >= 0 selects the register-resident
 * kernels (each lane holds its values in registers between LDS exchanges;
 * the tensor product happens in registers): rows have variant 0 for
 * Px/2 in {128, 256, 384, 512, 1024, 1536, 2048, 4096}; columns have
 * variant 0 (8 values per component and lane) for Py in {128, 256, 512,
 * 1024, 2048, 4096, 8192}, variant 1 (16 per lane) for Py >= 1024, and
 * variants 2, 3, 4 (variant 0's transforms on 2, 4 or 8 adjacent columns
 * per workgroup, the spectrum read and written in runs of that many
 * values; the same bits as variant 0) for Py in {512, 1024, 2048}, and
 * variants 5, 6 (variant 1's transforms, each on one wave, in workgroups
 * of 2 or 4 adjacent columns' x/y or 4 or 8 columns' z; the same bits as
 * variant 1) for Py = 1024.
 * -1 selects the LDS Stockham kernels (any 2^a 3^b 5^c length).  Default:
 * rows variant 0, columns variant 5 where it exists, else 3, else 0.
 * Returns -1 (nothing changed) when a requested variant does not exist for
 * the grid.  The families round differently (all are within 1e-12
 * relative of the exact convolution). */
int  Demag_SetFftKernels(DemagData* d, int row_variant, int col_variant);
int  Demag_GetFftKernels(DemagData* d, int* row_variant, int* col_variant);
/* Autotuner: times Demag_Apply (reps launches each, each after a pass
 * over a 384 MB scratch buffer that leaves the caches cold, as the
 * integrator's vector passes do) over the FFT kernel variants and, for the
 * Stockham kernels, the row/column workgroup sizes and LDS paddings; keeps
 * the fastest and returns it in choice[6] =
 * {row block, col block, row pad shift, col pad shift, row variant,
 * col variant}. */
int  Demag_Autotune(DemagData* d, int reps, int* choice);
/* introspection used by tests and the bench roofline */
int  Demag_GetPlanShape(DemagData* d, int* px, int* py, int* padded);
/* copy the (strength-scaled, 1/(px*py)-normalised) tensor spectrum to host:
 * planes xx, yy, zz, xy_re, xy_im, each py*(px/2+1) doubles, [ky][kx] */
int  Demag_GetSpectrum(DemagData* d, double* h_planes);
/* last error string of this module ("" when none) */
const char* Demag_LastError(void);

/* Newell tensor on the host: the nine components of `calt` for an
 * mdx x mdy displacement grid centred at (mdx/2, mdy/2), each plane
 * mdx*mdy doubles in the reference order aa,ab,ac,ba,bb,bc,ca,cb,cc
 * (2d_i4/demag_fft.cu:64-91).  nthreads<=0 uses all host cores. */
int  Demag_NewellTensor(double thick, int mdx, int mdy, double* planes9, int nthreads);

#ifdef __cplusplus
}
#endif
#endif /* LLG_DEMAG_FFT_H */
