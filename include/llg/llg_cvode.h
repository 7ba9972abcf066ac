/*
 * llg_cvode.h -- CVODE-compatible variable-order BDF integrator (Nordsieck
 * history, Newton iteration, SPGMR with left/right preconditioning) that
 * drives the LLG plug-ins when SUNDIALS is not installed.
 *
 * The reference calls SUNDIALS CVODE 7.x (third-party, not vendored;
 * src/irregular/2d_fft/2d_i4/Makefile:2-13) through
 *   CVodeCreate(CV_BDF) / CVodeInit / CVodeSetUserData / CVodeSVtolerances /
 *   SUNNonlinSol_Newton / SUNLinSol_SPGMR(PREC_LEFT, KRYLOV_DIM) /
 *   CVodeSetJacTimes / CVodeSetPreconditioner / SetGSType / SetMaxOrd / CVode
 * (2d_i4/2d_fft.cu:619-695).  This header offers the same sequence under an
 * LLGCV prefix, so it can coexist with a real SUNDIALS.  The algorithm
 * restates SUNDIALS' published CVODE/CVLS/SPGMR/Newton method and default
 * constants (no reference source exists for it); it works on any N_Vector
 * whose ops table is filled (HIP or serial).
 */
#ifndef LLG_CVODE_H
#define LLG_CVODE_H

#include "sundials_compat.h"

#ifdef __cplusplus
extern "C" {
#endif

#define LLGCV_BDF     2
#define LLGCV_NORMAL  1
#define LLGCV_ONE_STEP 2

/* return codes (CVODE values) */
#define LLGCV_SUCCESS          0
#define LLGCV_TSTOP_RETURN     1
#define LLGCV_WARNING          99
#define LLGCV_TOO_MUCH_WORK   -1
#define LLGCV_TOO_MUCH_ACC    -2
#define LLGCV_ERR_FAILURE     -3
#define LLGCV_CONV_FAILURE    -4
#define LLGCV_LINIT_FAIL      -5
#define LLGCV_LSETUP_FAIL     -6
#define LLGCV_LSOLVE_FAIL     -7
#define LLGCV_RHSFUNC_FAIL    -8
#define LLGCV_FIRST_RHSFUNC_ERR -9
#define LLGCV_REPTD_RHSFUNC_ERR -10
#define LLGCV_UNREC_RHSFUNC_ERR -11
#define LLGCV_MEM_FAIL        -20
#define LLGCV_MEM_NULL        -21
#define LLGCV_ILL_INPUT       -22
#define LLGCV_TOO_CLOSE       -27

typedef struct llgcv_stats {
  long nst, nfe, nsetups, nni, ncfn, netf, nli, nlcf, njvevals, npe, nps, nge;
  long nfeLS;
  int  qlast, qcur;
  double hlast, hcur, tcur, h0u;
  long nsyncs;            /* host synchronisations made by vector reductions */
} llgcv_stats;

/* Optional one-pass kernels for integrator-internal vector sequences.  Each
 * must produce exactly the values of the standard-op sequence it replaces:
 *   pascal(q, zn, sign): for k=1..q, j=q..k: zn[j-1] += sign*zn[j]
 *                        (the Nordsieck predict, sign=+1, and restore, -1)
 *   ewt(y, rtol, atol, abstol, w): w = 1/(rtol*|y| + atol) with the vector
 *                        abstol when non-NULL, else the scalar atol
 * and, inside SPGMR (each may be NULL; a non-zero return falls back to the
 * standard ops):
 *   scalediv(c, x, s, xout, vout): xout = c x; vout = xout / s
 *   proddotmulti(s, t, z, k, Y, d): z = s .* t; d_j = <z, Y_j>, j < k
 *   lincombdot(k, c, X, z, zz):   z = sum_j c_j X_j; *zz = <z, z>
 *   gsproject(s, t, z, k, Y, d, zz): z = s .* t; d_j = <z, Y_j> (Y_{k-1}
 *                        is z); z -= sum_{j<k-1} d_j Y_j; *zz = <z, z>
 *   lincombdiv(k, c, X, s, x): x = (0 + sum_j c_j X_j) / s
 * and in the Newton iteration:
 *   newtonupd(iter, ycor, del, zn0, w, y, dn, yn): ycor += del;
 *                        y = zn0 + ycor; *dn = WRMS(del, w) and, for
 *                        iter > 0, *yn = WRMS(ycor, w) computed as
 *                        N_VWrmsNormVectorArray does (iter 0: N_VWrmsNorm)
 *   atimes(v, z, gamma, t, y, fy, user_data): z = v - gamma J v in one
 *                        call, for use with the same jtimes callback (equal
 *                        to jtimes + N_VLinearSum(1, v, -gamma, Jv, z))
 * and, for left preconditioning with classical Gram-Schmidt, the whole
 * operator chain of one SPGMR iteration (both needed; used together):
 *   krylov(c, vl, s, vk, gamma, t, y, fy, user_data): with u = (c vl) / s,
 *                        vk = s .* psolve(u - gamma J u), vl left unscaled
 *                        (the registered jtimes and psolve callbacks, fused)
 *   gsscaled(z, k, Y, i, c, d, zz): Y_i = c Y_i, then the gsproject
 *                        projection of z (already formed) against Y      */
/* State a device-driven lagged SPGMR solve ends in (llgcv_fused_ops.lag_*):
 * status LLGCV_LAG_*, the Hessenberg columns formed, Krylov iterations
 * started, passes run after the first, |u_0|, the final residual
 * estimate, |u_l|, the Hessenberg matrix (rows 0..maxl) and the Givens
 * rotations (c, s) -- the values the host loop would hold. */
#define LLGCV_LAG_RUNNING (-1)
#define LLGCV_LAG_CONVERGED 0
#define LLGCV_LAG_MAXL 1
#define LLGCV_LAG_ZERO 2
#define LLGCV_LAG_BNORM 3
#define LLGCV_LAG_FALLBACK 4
#define LLGCV_LAG_QRFACT 5
#define LLGCV_LAG_DIM 16
typedef struct llgcv_lag_result {
  int status, krydim, iters, passes;
  sunrealtype r_norm, rho;
  sunrealtype nu[LLGCV_LAG_DIM];
  sunrealtype H[LLGCV_LAG_DIM][LLGCV_LAG_DIM];
  sunrealtype q[2 * LLGCV_LAG_DIM];
} llgcv_lag_result;

typedef struct llgcv_fused_ops {
  int (*pascal)(int q, N_Vector* zn, sunrealtype sign);
  int (*ewt)(N_Vector y, sunrealtype rtol, sunrealtype atol, N_Vector abstol, N_Vector w);
  int (*scalediv)(sunrealtype c, N_Vector x, N_Vector s, N_Vector xout, N_Vector vout);
  int (*proddotmulti)(N_Vector s, N_Vector t, N_Vector z, int k, N_Vector* Y, sunrealtype* d);
  int (*lincombdot)(int k, sunrealtype* c, N_Vector* X, N_Vector z, sunrealtype* zz);
  int (*gsproject)(N_Vector s, N_Vector t, N_Vector z, int k, N_Vector* Y, sunrealtype* d,
                   sunrealtype* zz);
  int (*lincombdiv)(int k, sunrealtype* c, N_Vector* X, N_Vector s, N_Vector x);
  int (*newtonupd)(int iter, N_Vector ycor, N_Vector del, N_Vector zn0, N_Vector w, N_Vector y,
                   sunrealtype* dn, sunrealtype* yn);
  int (*atimes)(N_Vector v, N_Vector z, sunrealtype gamma, sunrealtype t, N_Vector y,
                N_Vector fy, void* user_data);
  int (*krylov)(sunrealtype c, N_Vector vl, N_Vector s, N_Vector vk, sunrealtype gamma,
                sunrealtype t, N_Vector y, N_Vector fy, void* user_data);
  int (*gsscaled)(N_Vector z, int k, N_Vector* Y, int i, sunrealtype c, sunrealtype* d,
                  sunrealtype* zz);
  /* the whole left-preconditioned classical-GS SPGMR solve after the first
   * preconditioner solve (t0 = V[0] = P^-1 b), one host wait (status as
   * NVHIP_SPGMR_*; 5 = run the host-driven solve instead) */
  int (*spgmr)(N_Vector t0, N_Vector* V, int maxl, N_Vector s, N_Vector x, sunrealtype delta,
               sunrealtype gamma, sunrealtype t, N_Vector y, N_Vector fy, void* user_data,
               int* status, int* iters, sunrealtype* r_norm, sunrealtype* resnorm);
  /* SPGMR with lagged normalisation (left preconditioning, classical GS,
   * maxl <= 8): one pass and one host wait per iteration, the operator
   * applied to the unnormalised projection of the previous step (as
   * LLG_Arnoldi), the last projection norm by projnorm (as NVHip_ProjNorm).
   * Both must be set; any failure runs the standard solve instead. */
  int (*arnoldi)(int l, int pre, sunrealtype a, const sunrealtype* b, N_Vector win,
                 N_Vector* U, N_Vector wout, N_Vector s, sunrealtype gamma, sunrealtype t,
                 N_Vector y, N_Vector fy, void* user_data, sunrealtype* r);
  int (*projnorm)(int nb, sunrealtype a, const sunrealtype* b, N_Vector w, N_Vector* U,
                  N_Vector uout, sunrealtype* nn);
  /* the lagged-normalisation solve pipelined (all four needed): a pass
   * queued without waiting, with its coefficients and stop flag read from
   * device memory when given (as LLG_ArnoldiQueue), the wait for its sums
   * (as NVHip_ReduceWait), and a gate that holds the queued pass until the
   * host publishes its coefficients (as NVHip_GateOpen/Release) */
  int (*arnoldi_queue)(int l, int pre, sunrealtype a, const sunrealtype* b,
                       const double* coef_dev, const int* skip_dev, N_Vector win, N_Vector* U,
                       N_Vector wout, N_Vector s, sunrealtype gamma, sunrealtype t, N_Vector y,
                       N_Vector fy, void* user_data, unsigned* ticket);
  int (*result_wait)(N_Vector v, unsigned ticket, int K, sunrealtype* r);
  int (*gate_open)(N_Vector v, const double** coef_dev, const int** skip_dev);
  int (*gate_release)(N_Vector v, int stop, const sunrealtype* coef, int n);
  /* with arnoldi: the first pass with the right-hand side formed where it is
   * read, b = sum_k c_k X_k (as LLG_ArnoldiFirst; the Newton residual is then
   * never written), and the first Newton update with the solution formed
   * where it is read (as NVHip_NewtonUpdateComb; the solution vector is then
   * never written) */
  int (*arnoldi_first)(int nb, const sunrealtype* c, N_Vector* X, N_Vector* U, N_Vector wout,
                       N_Vector s, sunrealtype gamma, sunrealtype t, N_Vector y, N_Vector fy,
                       void* user_data, sunrealtype* r);
  int (*newtonupd_comb)(N_Vector ycor, int nvec, const sunrealtype* c, N_Vector* X,
                        N_Vector s, N_Vector zn0, N_Vector w, N_Vector y, sunrealtype* dnorm);
  /* predictor + acor = 0 + y = zn[0] + acor in one pass, with ewt not NULL
   * first the error weights from the old zn[0] (as NVHip_PredictInit) */
  int (*predict_init)(int q, N_Vector* zn, N_Vector acor, N_Vector y, sunrealtype rtol,
                      sunrealtype atol, N_Vector abstol, N_Vector ewt);
  /* nrm[0] = |x|_WRMS(w), nrm[1] = |a y + b z|_WRMS(w) (as NVHip_WrmsNormPair) */
  int (*wrms_pair)(N_Vector x, sunrealtype a, N_Vector y, sunrealtype b, N_Vector z,
                   N_Vector w, sunrealtype* nrm);
  /* the Newton residual res = c0 zn1 + c1 ycor + c2 f(t, y) with f formed in
   * the same pass and not stored (as LLG_RhsResidual: 0 done, 2 not handled,
   * then the RHS callback and the combination run; < 0 RHS failure).  f
   * is never stored, so the fy a later jtimes or preconditioner setup call
   * receives would be stale: the integrator takes this path only while the
   * registered jtimes and preconditioner setup are the two callbacks named
   * in fy_free_jtimes / fy_free_psetup (ones that never read fy; a NULL
   * psetup means none may be registered), and never with a jtimes setup. */
  int (*rhs_residual)(sunrealtype t, N_Vector y, const sunrealtype* c, N_Vector zn1,
                      N_Vector ycor, N_Vector res, void* user_data);
  CVLsJacTimesVecFn fy_free_jtimes;
  CVLsPrecSetupFn fy_free_psetup;
  /* with predict_init: the step's completion zn += l acor left to the next
   * predictor's pass (as NVHip_PredictComplete, with the rescale of a step
   * size change), and the order-change norm of the completed zn[q] taken
   * from the pending form (as NVHip_WrmsNormPairComp); any other reader of
   * zn gets the completion written first */
  int (*predict_complete)(int q, N_Vector* zn, N_Vector acor, const sunrealtype* l,
                          const sunrealtype* c, N_Vector y, sunrealtype rtol, sunrealtype atol,
                          N_Vector abstol, N_Vector ewt);
  int (*wrms_pair_comp)(N_Vector x, sunrealtype cx, N_Vector xc, sunrealtype a, N_Vector y,
                        sunrealtype b, N_Vector z, N_Vector w, sunrealtype* nrm);
  /* SPGMR with lagged normalisation driven by the device: with all five
   * set (and the lagged form in use, its first pass not formed in place),
   * the integrator queues a solve in at most two parts -- lag_begin, the
   * first pass and as many further passes as the previous solve needed
   * plus one by arnoldi_lag (as LLG_ArnoldiLag), the last with last = 1;
   * lag_wait; and when the state comes back running, lag_next, the
   * remaining passes and the last column's projection norm (lag_projnorm),
   * lag_wait -- for the state the host loop would reach; the scalar steps
   * run on the device (as NVHip_Lag*), so the solve is the host loop's bit
   * for bit. */
  int (*lag_begin)(N_Vector v, int maxl, sunrealtype delta, sunrealtype bnorm_floor,
                   sunrealtype nlen, const double** coef_dev, const int** skip_dev,
                   const int** pskip_dev);
  int (*arnoldi_lag)(int l, int pre, int last, const double* coef_dev, const int* skip_dev,
                     N_Vector win, N_Vector* U, N_Vector wout, N_Vector s, sunrealtype gamma,
                     sunrealtype t, N_Vector y, N_Vector fy, void* user_data);
  int (*lag_projnorm)(N_Vector w, int nb, N_Vector* U);
  int (*lag_wait)(N_Vector v, llgcv_lag_result* res);
  int (*lag_next)(N_Vector v);
  /* with the device-driven solve, on a Newton iteration's first solve:
   * the solution and the first Newton update queued behind each part of
   * the solve before its wait (as NVHip_LagUpdate; they do nothing unless
   * the solve ends there with a solution the iteration uses), and
   * lag_update_wait for the update's norm (as NVHip_LagUpdateWait) -- one
   * host wait less per Newton iteration, the same bits */
  int (*lag_update)(N_Vector b, int nvec, N_Vector* V, N_Vector s, N_Vector x, N_Vector ycor,
                    N_Vector zn0, N_Vector w, N_Vector y, unsigned* ticket);
  int (*lag_update_wait)(N_Vector ycor, unsigned ticket, sunrealtype* dnorm);
  /* with lag_update, on a step that will test an order change when it
   * succeeds: the test's two norms (wrms_pair_comp's, from the pending
   * completion and the step's coefficients) queued behind the update
   * (as NVHip_WrmsNormPairCompQueue) and read after the step (as
   * NVHip_WrmsNormPairWait) when the first Newton iteration converged --
   * one host wait less on such a step, the same bits */
  int (*pair_queue)(N_Vector x, sunrealtype cx, N_Vector xc, sunrealtype a, N_Vector y,
                    sunrealtype b, N_Vector z, N_Vector w, unsigned* ticket);
  int (*pair_wait)(N_Vector x, unsigned ticket, sunrealtype* nrm);
  /* with predict_complete: the copy of acor into zn[qmax] that a completed
   * step keeps for the next order-change test left to the next predictor's
   * pass too (as NVHip_PredictCompleteSave) */
  int (*predict_complete_save)(int q, N_Vector* zn, N_Vector acor, const sunrealtype* l,
                               const sunrealtype* c, N_Vector y, sunrealtype rtol,
                               sunrealtype atol, N_Vector abstol, N_Vector ewt, N_Vector save);
} llgcv_fused_ops;

void* LLGCVodeCreate(int lmm, SUNContext sunctx);
int   LLGCVodeSetFusedOps(void* mem, const llgcv_fused_ops* ops);
int   LLGCVodeInit(void* mem, CVRhsFn f, sunrealtype t0, N_Vector y0);
int   LLGCVodeReInit(void* mem, sunrealtype t0, N_Vector y0);
int   LLGCVodeSetUserData(void* mem, void* user_data);
int   LLGCVodeSStolerances(void* mem, sunrealtype rtol, sunrealtype atol);
int   LLGCVodeSVtolerances(void* mem, sunrealtype rtol, N_Vector abstol);
int   LLGCVodeSetMaxOrd(void* mem, int maxord);
int   LLGCVodeSetMaxNumSteps(void* mem, long mxsteps);
int   LLGCVodeSetInitStep(void* mem, sunrealtype hin);
int   LLGCVodeSetMaxStep(void* mem, sunrealtype hmax);
int   LLGCVodeSetMinStep(void* mem, sunrealtype hmin);
int   LLGCVodeSetStopTime(void* mem, sunrealtype tstop);
/* SPGMR linear solver: pretype SUN_PREC_*, maxl Krylov dim (<=0 -> 5),
 * gstype SUN_MODIFIED_GS / SUN_CLASSICAL_GS */
int   LLGCVodeSetLinearSolverSPGMR(void* mem, int pretype, int maxl, int gstype);
int   LLGCVodeSetEpsLin(void* mem, sunrealtype eplifac);
int   LLGCVodeSetJacTimes(void* mem, CVLsJacTimesSetupFn jtsetup, CVLsJacTimesVecFn jtimes);
int   LLGCVodeSetPreconditioner(void* mem, CVLsPrecSetupFn pset, CVLsPrecSolveFn psolve);
int   LLGCVode(void* mem, sunrealtype tout, N_Vector yout, sunrealtype* tret, int itask);
int   LLGCVodeGetDky(void* mem, sunrealtype t, int k, N_Vector dky);
int   LLGCVodeGetStats(void* mem, llgcv_stats* st);
void  LLGCVodeFree(void** mem);

#ifdef __cplusplus
}
#endif
#endif /* LLG_CVODE_H */
