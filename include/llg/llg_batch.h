/*
 * llg_batch.h -- one launch per pass over many LLG instances.
 *
 * No reference counterpart: the reference evaluates one instance per
 * process (its f() wrapper src/irregular/2d_fft/2d_i4/2d_fft.cu:328-366 and
 * JtvProduct 2d_i4/jtv.cu:47-147, one grid each).  A batch handle holds a
 * device table of its instances' fixed data (geometry, term sets, demag
 * field buffer); a call then evaluates f (or Jv) for any subset of them:
 *   - unit-sphere normalisation: one launch over all instances;
 *   - demag: the three register-FFT passes, one launch each per group of
 *     instances that share an FFT plan (grid shape, padding, kernel
 *     variants), blockIdx.y selecting the instance; instances without a
 *     register plan (or in the FP32 spectral form) run their own passes;
 *   - the RHS (or Jv) stencil: one launch over all instances of any grid
 *     shapes, each workgroup finding its instance and tile in a prefix of
 *     the instances' tile counts (LDS-tiled, 32x8 cells per tile).
 * Every instance gets exactly the bits of its own LLG_Rhs / JtvProduct
 * call: the same per-cell (and per-transform) arithmetic runs for it.
 * Dense-mask (LLG_EXEC_YMSK) instances only.  A handle is used from one
 * stream at a time; its launches go to the stream passed to the call.
 */
#ifndef LLG_BATCH_H
#define LLG_BATCH_H

#include "llg_plugin.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The Newton residual formed in the RHS pass instead of f (LLG_RhsResidual):
 * out = c0 zn1 + c1 ycor + c2 f, device arrays of the instance's size. */
typedef struct llg_res_in {
  const double* zn1;
  const double* ycor;
  double c0, c1, c2;
} llg_res_in;

typedef struct llg_batch llg_batch;

/* n >= 1 contexts (dense-mask execution, one device).  NULL on failure
 * (LLG_LastError). */
llg_batch* LLG_BatchCreate(int n, llg_ctx* const* ctx);
int        LLG_BatchSize(const llg_batch* b);

/* f for every instance i with y[i] != NULL: ydot[i] = f(y[i]) (y[i] is
 * normalised in place when the instance's f normalises), or, when res and
 * res[i] are non-NULL, the Newton residual of res[i] (LLG_RhsResidual).
 * Device arrays; the demag field of instance i lands in its context's
 * field buffer (read by PrecondSetup, as after LLG_Rhs).  Returns 0, or -1
 * on a launch error. */
int LLG_RhsBatch(llg_batch* b, double* const* y, double* const* ydot,
                 const llg_res_in* const* res, void* stream);

/* Jv[i] = J(y[i]) v[i] (JtvProduct) for every instance i with v[i] != NULL. */
int LLG_JtvBatch(llg_batch* b, const double* const* v, double* const* Jv,
                 const double* const* y, void* stream);

void LLG_BatchDestroy(llg_batch* b);

#ifdef __cplusplus
}
#endif
#endif /* LLG_BATCH_H */
