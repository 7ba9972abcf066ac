/*
 * llg_driver.h -- the reference's main() time loop as a library call.
 *
 * Replaces the setup and output loop of src/irregular/2d_fft/2d_i4/2d_fft.cu
 * :545-700 (N_VNew x2, FusedNVec_Init, CVodeInit/SVtolerances/SPGMR/
 * SetJacTimes/SetPreconditioner/SetMaxOrd, then CVode every T1 = 0.1 inside
 * a timed region) and PrintFinalStats (:369-390, via llgcv_stats).
 */
#ifndef LLG_DRIVER_H
#define LLG_DRIVER_H

#include "llg_cvode.h"
#include "llg_plugin.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct llg_run_opts {
  double t0;          /* initial time */
  double t_out;       /* output cadence T1 */
  double rtol, atol;  /* RTOL_VAL / ATOL_VAL */
  int    maxl;        /* KRYLOV_DIM */
  int    gstype;      /* 0 = auto (classical iff neq < 500000), else SUN_*_GS */
  int    max_order;   /* MAX_BDF_ORDER */
  int    fused;       /* bit 0: FusedNVec_Init on y (one-launch multi-dot and
                         linear combination); bit 1: the integrator's one-pass
                         kernels (Nordsieck, error weights, SPGMR products,
                         Gram-Schmidt); bit 2 (with bit 1, left block-Jacobi
                         and analytic Jv): each SPGMR iteration's operator
                         chain in one launch (LLG_Krylov); bit 3 (with bits
                         1 and 2, classical GS): the whole SPGMR solve
                         queued on the device with one host wait
                         (LLG_SpgmrSolve; it halves the host waits but
                         queues every iteration up to the Krylov
                         dimension, which costs more than the waits on
                         the 1536x512 workload); bit 4 (with bits 1 and 2,
                         classical GS, maxl <= 8, dense grid): SPGMR with
                         lagged normalisation, one pass and one host wait
                         per iteration (LLG_Arnoldi; same method, other
                         rounding); bit 5 (with bit 4, sync mode 1 or 2):
                         each pass of that solve queued behind a gate
                         (NVHip_GateOpen) before the previous pass's sums
                         are back, taking the launch off the host round
                         trip (on the 1536x512 workload it measures level
                         with bit 4 alone: the gate adds a kernel boundary
                         per iteration); bits 6 and 7 (with bit 4): the
                         first SPGMR pass forms the Newton residual, and
                         the first Newton update the SPGMR solution, where
                         each is read (never written; measured neutral on
                         the 1536x512 workload); bit 8 (with bit 1 and
                         analytic Jv): the Newton residual formed in the
                         RHS pass (LLG_RhsResidual; f is not stored);
                         bit 9 (with bit 1): each step's completion
                         zn += l acor left to the next predictor's pass
                         (NVHip_PredictComplete); bit 10 (with bit 4, its
                         first pass not formed in place): each lagged
                         SPGMR solve queued whole, its scalar steps on
                         the device (NVHip_Lag*), one host wait per
                         solve; bit 11 (with bit 10): the solution and
                         the first Newton update of each iteration queued
                         behind the solve (NVHip_LagUpdate), no host
                         round trip between them; bit 12 (with bit 11):
                         on a step that will test an order change, the
                         test's two norms queued behind the first Newton
                         update (NVHip_WrmsNormPairCompQueue), one host
                         round trip less on such a step; bit 13 (with bit
                         9): a completed step's copy of acor into
                         zn[qmax] left to the next predictor's pass too
                         (NVHip_PredictCompleteSave); default 16151 (bits
                         0, 1, 2, 4, 8 to 13).  Bits 1 to 3 and 8 to 13
                         change no value (bits 10 to 12 none of bit 4's). */
  long   mxstep;      /* max steps per CVode call (CVODE default 500) */
  int    use_jtv;     /* analytic Jv (else difference quotient) */
  int    use_prec;    /* block-Jacobi preconditioner */
  int    pretype;     /* SUN_PREC_LEFT (reference) or SUN_PREC_RIGHT */
  int    pad;
} llg_run_opts;

typedef struct llg_run_result {
  llgcv_stats st;
  double wall_s;        /* host wall time of the output loop */
  double gpu_ms;        /* hipEvent time of the same loop */
  double sim_per_wall;  /* simulated time advanced / wall_s */
  double t_reached;
  long   nout;
  int    retval;
  int    pad;
} llg_run_result;

void  LLG_DefaultRunOpts(llg_run_opts* o);
void* LLG_RunCreate(llg_ctx* ctx, const llg_run_opts* opts, const double* y0_host);
int   LLG_RunAdvance(void* run, double t_end, llg_run_result* res);
int   LLG_RunGetState(void* run, double* y_host);
/* One quasi-static stage of a field sweep: set the applied field of the
 * term sets in `which` (LLG_SetAppliedField), restart the integrator at
 * the run's current time and state (a field step is a discontinuity of f,
 * so the BDF history is dropped, as CVodeReInit does), advance by t_stage
 * on the output cadence, and return the stage's statistics in res and the
 * mean magnetisation of the active cells at its end in m_avg[3].  No
 * reference counterpart: the reference's field is a compile-time constant
 * (src/irregular/2d_fft/2d_i4/2d_fft.cu:178-180). */
int   LLG_RunStage(void* run, int which, const double* happ, double t_stage,
                   llg_run_result* res, double* m_avg);
/* A field ramp of nstages stages (LLG_RunStage each): stage s applies
 * happ3[3s .. 3s+2] for t_stage; m_avg[3s .. 3s+2] receives its final mean
 * magnetisation.  res sums the counters and wall time of all stages
 * (t_reached is the last stage's).  Returns 0, or the first failing stage's
 * negative code (stages after it are not run). */
int   LLG_RunRamp(void* run, int which, int nstages, const double* happ3, double t_stage,
                  double* m_avg, llg_run_result* res);
void  LLG_RunDestroy(void* run);

#ifdef __cplusplus
}
#endif
#endif /* LLG_DRIVER_H */
