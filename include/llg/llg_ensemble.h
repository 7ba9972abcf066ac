/*
 * llg_ensemble.h -- many independent LLG instances (different grids,
 * geometries, fields) integrated concurrently on one GPU, each with its own
 * BDF step control.
 *
 * No reference counterpart: the reference runs one instance per process
 * (src/irregular/2d_fft/2d_i4/2d_fft.cu main()).  Each instance here is the
 * reference's whole solver -- its own context (LLG_Create), HIP stream,
 * N_Vectors and integrator (LLG_RunCreate) -- and one host thread drives
 * each instance, so the GPU overlaps the instances' kernels and hides each
 * one's host round trips behind the others' work.  An instance computes
 * exactly what it computes alone (same kernels, same launch shapes, same
 * reduction order): results are bitwise those of single-instance runs.
 * Optionally (LLG_EnsembleSetBatched) the f evaluations of same-shape
 * instances are gathered into one LLG_RhsBatch each (llg_batch.h), with
 * the same bits.
 */
#ifndef LLG_ENSEMBLE_H
#define LLG_ENSEMBLE_H

#include "llg_driver.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct llg_instance {
  llg_config cfg;               /* grid, physics and demag of this instance */
  const unsigned char* active;  /* ng*ny active flags, NULL = all active */
  const double* w;              /* per-cell window (GATE_WINDOW) or NULL */
  const double* cell_msk;       /* per-cell axes (cell_axes) or NULL */
  const double* cell_nsk;
  const double* y0;             /* 3*ng*ny initial state (host) */
} llg_instance;

typedef struct llg_ensemble_result {
  double wall_s;       /* wall time of the advance */
  double inst_time;    /* simulated time advanced, summed over instances */
  long   nst, nfe, nli, nni;
  int    failed;       /* instances whose integration returned an error */
  int    pad;
} llg_ensemble_result;

/* n instances, each with the run options `opts`; nthreads host threads
 * (<= 0: one per instance).  NULL on failure (LLG_LastError). */
void* LLG_EnsembleCreate(int n, const llg_instance* inst, const llg_run_opts* opts,
                         int nthreads);
/* advance every instance to t_end (each on its own output cadence) */
int   LLG_EnsembleAdvance(void* ens, double t_end, llg_ensemble_result* res);
int   LLG_EnsembleSize(void* ens);
int   LLG_EnsembleGetState(void* ens, int i, double* y_host);
/* statistics of instance i's last advance */
int   LLG_EnsembleGetResult(void* ens, int i, llg_run_result* r);
/* mode 1: instances of the same grid shape evaluate f together -- each
 * instance's f() call waits until every live instance of its shape has
 * called f, and one LLG_RhsBatch (llg_batch.h) then evaluates all of them:
 * one normalisation launch, one launch per demag pass and one RHS launch
 * for the group instead of one set per instance.  Each instance keeps its
 * own step control, and its results stay bitwise those of its own run.
 * Every instance then runs on a host thread of its own.  Dense-mask
 * instances only.  mode 2 batches only the shapes of at most 2^17 cells
 * (whose instances are bound by kernel dispatch, not memory traffic); the
 * others evaluate f on their own.  mode 0 restores independent
 * evaluation.  Returns the number of batched shape groups, or -1. */
int   LLG_EnsembleSetBatched(void* ens, int mode);
/* batched launches and the f() calls they served, summed over groups */
int   LLG_EnsembleBatchStats(void* ens, long* launches, long* calls);
void  LLG_EnsembleDestroy(void* ens);

#ifdef __cplusplus
}
#endif
#endif /* LLG_ENSEMBLE_H */
