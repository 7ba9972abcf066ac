/*
 * nvector_hip.h -- HIP N_Vector for gfx950 (MI355X): the device vector the
 * LLG plug-ins and the integrator operate on, plus the fused-ops layer.
 *
 * Replaces, on the reference's hot path:
 *   - SUNDIALS NVECTOR_CUDA as used by every variant (N_VNew_Cuda at
 *     src/irregular/2d_fft/2d_i4/2d_fft.cu:545-546, device/host pointer
 *     accessors at :332-333 and :554-555, copies at :612-613, :703);
 *   - the fused N_Vector layer FusedNVec_Init / FusedNVec_FreePool
 *     (2d_i4/deferred_nvector.h:58,66; multi-dot override
 *     2d_i4/deferred_nvector.cu:95-180; linear-combination override
 *     src/2d/2d_heter/deferred_nvector.cu:135-151,209-236; clone
 *     propagation 2d_i4/deferred_nvector.cu:182-202).
 *
 * Layout: one contiguous FP64 device array per vector (SoA [mx|my|mz] for
 * the LLG state), an optional host mirror, and the HIP stream every op of
 * the vector is launched on.  Reductions are deterministic: per-workgroup
 * partials in fixed slots, summed in slot order by a one-workgroup finish
 * launch that writes the result to pinned host memory (one host wait).
 */
#ifndef LLG_NVECTOR_HIP_H
#define LLG_NVECTOR_HIP_H

#include "sundials_compat.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Largest number of vectors a fused op handles in one launch; beyond it
 * the op loops over groups (the reference falls back to per-vector ops
 * above 16, 2d_i4/deferred_nvector.cu:151-155). */
#define NVHIP_FUSED_MAX 16

/* constructors -- stream==NULL selects the library's default stream */
N_Vector N_VNew_Hip(sunindextype length, SUNContext sunctx);
N_Vector N_VNewWithStream_Hip(sunindextype length, void* stream, SUNContext sunctx);
N_Vector N_VNewEmpty_Hip(SUNContext sunctx);
N_Vector N_VMake_Hip(sunindextype length, sunrealtype* d_data, void* stream,
                     SUNContext sunctx);

/* accessors (NVECTOR_CUDA equivalents) */
sunrealtype* N_VGetDeviceArrayPointer_Hip(N_Vector v);
sunrealtype* N_VGetHostArrayPointer_Hip(N_Vector v);
void         N_VCopyToDevice_Hip(N_Vector v);
void         N_VCopyFromDevice_Hip(N_Vector v);
void*        N_VGetStream_Hip(N_Vector v);
sunindextype N_VGetLength_Hip(N_Vector v);
int          N_VEnableFusedOps_Hip(N_Vector v, sunbooleantype tf);
/* host <-> device copies of raw arrays of the vector's length */
int          N_VSetFromHost_Hip(N_Vector v, const sunrealtype* h_src);
int          N_VGetToHost_Hip(N_Vector v, sunrealtype* h_dst);

/* the default stream used when a constructor gets stream==NULL */
void* NVHip_DefaultStream(void);
/* counters of host synchronisations made by reductions (diagnostics) */
long  NVHip_NumSyncs(void);
/* how the host waits for a reduction result: 0 = hipStreamSynchronize,
 * 1 = poll the pinned result's sequence word (bounded, falls back to 0),
 * 2 = as 1 but yielding the core between polls (many host threads) */
int   NVHip_SetSyncMode(int mode);
/* reduction launch shape: at most max_blocks (1..4096) partial workgroups;
 * mode 0 = one launch (last workgroup sums the partials), 1 = partial
 * launch + one-workgroup finish launch (default: 1024, 1).  Results are
 * deterministic for a fixed configuration and vector length. */
int   NVHip_SetReduceConfig(int max_blocks, int mode);

/* standard ops */
N_Vector_ID  N_VGetVectorID_Hip(N_Vector v);
N_Vector     N_VClone_Hip(N_Vector w);
N_Vector     N_VCloneEmpty_Hip(N_Vector w);
void         N_VDestroy_Hip(N_Vector v);
void         N_VSpace_Hip(N_Vector v, sunindextype* lrw, sunindextype* liw);
void         N_VLinearSum_Hip(sunrealtype a, N_Vector x, sunrealtype b, N_Vector y, N_Vector z);
void         N_VConst_Hip(sunrealtype c, N_Vector z);
void         N_VProd_Hip(N_Vector x, N_Vector y, N_Vector z);
void         N_VDiv_Hip(N_Vector x, N_Vector y, N_Vector z);
void         N_VScale_Hip(sunrealtype c, N_Vector x, N_Vector z);
void         N_VAbs_Hip(N_Vector x, N_Vector z);
void         N_VInv_Hip(N_Vector x, N_Vector z);
void         N_VAddConst_Hip(N_Vector x, sunrealtype b, N_Vector z);
sunrealtype  N_VDotProd_Hip(N_Vector x, N_Vector y);
sunrealtype  N_VMaxNorm_Hip(N_Vector x);
sunrealtype  N_VWrmsNorm_Hip(N_Vector x, N_Vector w);
sunrealtype  N_VWrmsNormMask_Hip(N_Vector x, N_Vector w, N_Vector id);
sunrealtype  N_VMin_Hip(N_Vector x);
sunrealtype  N_VWL2Norm_Hip(N_Vector x, N_Vector w);
sunrealtype  N_VL1Norm_Hip(N_Vector x);
void         N_VCompare_Hip(sunrealtype c, N_Vector x, N_Vector z);
sunbooleantype N_VInvTest_Hip(N_Vector x, N_Vector z);
sunbooleantype N_VConstrMask_Hip(N_Vector c, N_Vector x, N_Vector m);
sunrealtype  N_VMinQuotient_Hip(N_Vector num, N_Vector denom);
sunrealtype  N_VWSqrSumLocal_Hip(N_Vector x, N_Vector w);
sunrealtype  N_VWSqrSumMaskLocal_Hip(N_Vector x, N_Vector w, N_Vector id);

/* fused ops */
SUNErrCode N_VLinearCombination_Hip(int nvec, sunrealtype* c, N_Vector* X, N_Vector z);
SUNErrCode N_VScaleAddMulti_Hip(int nvec, sunrealtype* a, N_Vector x, N_Vector* Y, N_Vector* Z);
SUNErrCode N_VDotProdMulti_Hip(int nvec, N_Vector x, N_Vector* Y, sunrealtype* dotprods);
/* N_VDotProdMulti_Hip without any allocation: no counterpart in the
 * reference.  For callers that cannot allocate device or pinned memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The partial sums live in a file-scope device array of
 * 1024 x 16 + 17 doubles (131 KB) that each GPU holds until the process
 * ends; the results come back by a device-to-host copy into `dotprods` and
 * one stream synchronisation.  nvec must be 1..16 (NVHIP_FUSED_MAX; -1 and
 * nothing launched otherwise, or on vectors of unequal length).  Every call
 * on one GPU shares that array, so calls that overlap on one GPU (two host
 * threads, two streams) race.  With the library's default reduction shape
 * (at most 1024 partial workgroups) the results are bitwise those of
 * N_VDotProdMulti_Hip for nvec >= 2 (nvec = 1 takes the multi-dot kernel,
 * where N_VDotProdMulti_Hip takes the single-dot one). */
SUNErrCode N_VDotProdMultiNoAlloc_Hip(int nvec, N_Vector x, N_Vector* Y, sunrealtype* dotprods);

/* vector-array ops */
SUNErrCode N_VLinearSumVectorArray_Hip(int nvec, sunrealtype a, N_Vector* X, sunrealtype b,
                                       N_Vector* Y, N_Vector* Z);
SUNErrCode N_VScaleVectorArray_Hip(int nvec, sunrealtype* c, N_Vector* X, N_Vector* Z);
SUNErrCode N_VConstVectorArray_Hip(int nvec, sunrealtype c, N_Vector* Z);
SUNErrCode N_VWrmsNormVectorArray_Hip(int nvec, N_Vector* X, N_Vector* W, sunrealtype* nrm);
SUNErrCode N_VWrmsNormMaskVectorArray_Hip(int nvec, N_Vector* X, N_Vector* W, N_Vector id,
                                          sunrealtype* nrm);

/* One-pass kernels for the integrator's vector sequences (llgcv_fused_ops):
 * the Nordsieck predict/restore over q+1 history vectors and the error
 * weights.  Values are identical to the equivalent N_VLinearSum / N_VAbs /
 * N_VInv sequences. */
int NVHip_Pascal(int q, N_Vector* zn, sunrealtype sign);
int NVHip_EwtSet(N_Vector y, sunrealtype rtol, sunrealtype atol, N_Vector abstol, N_Vector w);
/* The predictor (NVHip_Pascal with sign +1) and the start of the Newton
 * iteration in one pass: acor = 0 (N_VConst) and y = zn[0] + acor
 * (N_VLinearSum(1, zn[0], 1, acor, y)); with ewt not NULL, first the error
 * weights of the step from the zn[0] being replaced (NVHip_EwtSet(zn[0],
 * rtol, atol, abstol, ewt)).  Bit-equal to those calls.  acor, y and ewt
 * may not alias zn or each other. */
int NVHip_PredictInit(int q, N_Vector* zn, N_Vector acor, N_Vector y, sunrealtype rtol,
                      sunrealtype atol, N_Vector abstol, N_Vector ewt);
/* NVHip_PredictInit preceded, in the same pass, by the previous step's
 * completion zn[j] += l[j] acor, j = 0..q (N_VScaleAddMulti(q+1, l, acor,
 * zn, zn)) and, with c not NULL, the step-size rescale zn[j] = c[j] zn[j],
 * j = 1..q (N_VScaleVectorArray; c[0] unused); the error weights (ewt not
 * NULL) come from the completed zn[0].  Bit-equal to those calls. */
int NVHip_PredictComplete(int q, N_Vector* zn, N_Vector acor, const sunrealtype* l,
                          const sunrealtype* c, N_Vector y, sunrealtype rtol, sunrealtype atol,
                          N_Vector abstol, N_Vector ewt);
/* NVHip_PredictComplete that also keeps the completed step's correction:
 * save = acor (N_VScale(1, acor, save)) before acor is zeroed, in the same
 * pass; save may not alias zn[0..q], acor, y or ewt.  No reference
 * counterpart. */
int NVHip_PredictCompleteSave(int q, N_Vector* zn, N_Vector acor, const sunrealtype* l,
                              const sunrealtype* c, N_Vector y, sunrealtype rtol,
                              sunrealtype atol, N_Vector abstol, N_Vector ewt, N_Vector save);
/* Two WRMS norms with one host wait: nrm[0] = |x|_WRMS(w) and nrm[1] =
 * |a y + b z|_WRMS(w), the combination not written; each equal to
 * N_VWrmsNorm (after N_VLinearSum for the second) bit for bit. */
int NVHip_WrmsNormPair(N_Vector x, sunrealtype a, N_Vector y, sunrealtype b, N_Vector z,
                       N_Vector w, sunrealtype* nrm);
/* NVHip_WrmsNormPair with nrm[0] the norm of x + cx xc as N_VScaleAddMulti
 * forms it (a pending completion of x), neither written. */
int NVHip_WrmsNormPairComp(N_Vector x, sunrealtype cx, N_Vector xc, sunrealtype a, N_Vector y,
                           sunrealtype b, N_Vector z, N_Vector w, sunrealtype* nrm);
/* NVHip_WrmsNormPairComp queued on x's stream without a wait: the result
 * lands in the second result slot (the one a multi-dot projection uses), so
 * reductions waited for on the first slot in between leave it alone; a
 * later queue or projection overwrites it.  NVHip_WrmsNormPairWait waits
 * for the pair queued as ticket and returns -1 when the slot holds another
 * result.  No reference counterpart. */
int NVHip_WrmsNormPairCompQueue(N_Vector x, sunrealtype cx, N_Vector xc, sunrealtype a,
                                N_Vector y, sunrealtype b, N_Vector z, N_Vector w,
                                unsigned* ticket);
int NVHip_WrmsNormPairWait(N_Vector x, unsigned ticket, sunrealtype* nrm);

/* One-pass kernels for the SPGMR loop (llgcv_fused_ops), each equal
 * element for element, and reduction for reduction, to the N_Vector
 * sequence it replaces:
 *   NVHip_ScaleDiv     xout = c x; vout = xout / s
 *                      (N_VScale then N_VDiv; vout may not alias x or s)
 *   NVHip_ProdDotMulti z = s .* t; dots_j = <z, Y_j>, j < nvec <= 16,
 *                      where Y_j may be z itself
 *                      (N_VProd then N_VDotProdMulti)
 *   NVHip_LinCombDot   z = sum_j c_j X_j (X_j may alias z); *zz = <z, z>
 *                      (N_VLinearCombination then N_VDotProd), nvec <= 16
 *   NVHip_GSProject    one classical Gram-Schmidt projection with ONE host
 *                      wait: z = s .* t; dots_j = <z, Y_j> (j < nvec, and
 *                      Y_{nvec-1} must be z); z -= sum_{j<nvec-1} dots_j Y_j;
 *                      *zz = <z, z>.  The combination reads its coefficients
 *                      from device memory, so it is queued behind the
 *                      multi-dot without a round trip.  Equal to N_VProd,
 *                      N_VDotProdMulti, N_VLinearCombination, N_VDotProd. */
int NVHip_ScaleDiv(sunrealtype c, N_Vector x, N_Vector s, N_Vector xout, N_Vector vout);
int NVHip_ProdDotMulti(N_Vector s, N_Vector t, N_Vector z, int nvec, N_Vector* Y,
                       sunrealtype* dots);
int NVHip_LinCombDot(int nvec, sunrealtype* c, N_Vector* X, N_Vector z, sunrealtype* zz);
int NVHip_GSProject(N_Vector s, N_Vector t, N_Vector z, int nvec, N_Vector* Y,
                    sunrealtype* dots, sunrealtype* zz);
/* NVHip_GSProject for a z already formed, with the basis vector Y_sidx
 * (sidx < nvec - 1) normalised in the same pass: Y_sidx = sc Y_sidx is
 * written back as the multi-dot reads it.  Equal to N_VScale(sc, Y_sidx)
 * followed by the projection of z. */
int NVHip_GSProjectScaled(N_Vector z, int nvec, N_Vector* Y, int sidx, sunrealtype sc,
                          sunrealtype* dots, sunrealtype* zz);
/* x = (0 + sum_j c_j X_j) / s (N_VConst 0, N_VLinearCombination, N_VDiv) */
int NVHip_LinCombDiv(int nvec, sunrealtype* c, N_Vector* X, N_Vector s, N_Vector x);
/* Newton update: ycor += del; y = zn0 + ycor; *dnorm = WRMS(del, w); for
 * iter > 0 also *ynorm = WRMS(ycor, w) (as N_VWrmsNormVectorArray of the
 * pair; iter 0 as N_VWrmsNorm).  y must not alias the inputs. */
int NVHip_NewtonUpdate(int iter, N_Vector ycor, N_Vector del, N_Vector zn0, N_Vector w,
                       N_Vector y, sunrealtype* dnorm, sunrealtype* ynorm);
/* The first Newton update with the SPGMR solution formed where it is read:
 * del = (0 + sum_j c_j X_j) / s (NVHip_LinCombDiv's values, not written),
 * then NVHip_NewtonUpdate(0, ycor, del, zn0, w, y, dnorm); bit-equal to the
 * two calls.  y may not alias any input, ycor no X_j. */
int NVHip_NewtonUpdateComb(N_Vector ycor, int nvec, const sunrealtype* c, N_Vector* X,
                           N_Vector s, N_Vector zn0, N_Vector w, N_Vector y,
                           sunrealtype* dnorm);

/* Device-resident SPGMR solve: SUNLinSolSolve_SPGMR with left
 * preconditioning, zero initial guess, classical Gram-Schmidt and no
 * restarts, queued whole on the stream of x with ONE host wait.  The
 * Hessenberg entries, Givens rotations, residual estimate and convergence
 * test live on the device; kernels of iterations after convergence exit at
 * once.  Same operations in the same order as the host-driven solve, so
 * the same bits.  No counterpart in the reference (its SUNDIALS solve waits
 * for every Gram-Schmidt reduction).
 *   op(c_dev, skip_dev, vl, s, vk, op_data) queues vk = s .* P^-1 (I -
 *     gamma J) ((*c_dev) vl / s) (the operator chain with the basis scale
 *     read from device memory); it must queue nothing that runs when
 *     *skip_dev is nonzero and return 0, or nonzero when it cannot run.
 *   t0 = V[0] holds P^-1 b on entry; V[0..maxl] are workspace
 *     (maxl <= NVHIP_FUSED_MAX - 1); sv the scaling (weight) vector; x the
 *     solution (may alias b, which is not read).
 *   *status: one of NVHIP_SPGMR_*; FALLBACK means the classical GS step
 *     needed a re-orthogonalisation pass: x is untouched and the caller
 *     runs its host-driven solve.  *iters Krylov iterations, *r_norm the
 *     initial residual norm, *resnorm the final residual estimate.
 * Returns 0, or -1 (nothing queued that changes x) when the arguments or
 * the reduction mode do not allow the device-resident form. */
#define NVHIP_SPGMR_SUCCESS 0
#define NVHIP_SPGMR_RES_REDUCED 1
#define NVHIP_SPGMR_CONV_FAIL 2
#define NVHIP_SPGMR_QRFACT_FAIL 3
#define NVHIP_SPGMR_QRSOL_FAIL 4
#define NVHIP_SPGMR_FALLBACK 5
typedef int (*NVHipDevOp)(const double* c_dev, const int* skip_dev, N_Vector vl, N_Vector s,
                          N_Vector vk, void* op_data);
int NVHip_SpgmrSolve(NVHipDevOp op, void* op_data, N_Vector t0, N_Vector* V, int maxl,
                     N_Vector sv, N_Vector x, sunrealtype delta, int* status, int* iters,
                     sunrealtype* r_norm, sunrealtype* resnorm);

/* Device-driven SPGMR with lagged normalisation (no counterpart in the
 * reference): the scalar step of each Krylov iteration of the integrator's
 * lagged SPGMR -- the Hessenberg column from a pass's sums, the Pythagorean
 * estimate of h_{l+1,l} and its trial rotation, the re-orthogonalisation
 * test, the Givens rotation, the residual estimate, the convergence test
 * and the next pass's projection coefficients -- runs on the device, on
 * thread 0 of the kernel that sums the pass's partials, with the host
 * loop's operations in its order.  A solve is queued whole: pass 0 (the
 * preconditioned right-hand side) with NVHip_LagStep(p = 0), passes
 * p = 1 .. maxl-1 reading their coefficients from *coef_dev and skipped
 * when *skip_dev is set, each with NVHip_LagStep(p), and the projection
 * norm of the last column (NVHip_LagProjNorm, skipped unless needed); the
 * host waits once (NVHip_LagWait) for the state the solve ends in.
 *   NVHip_LagBegin: v's stream; maxl <= 8; delta the residual tolerance;
 *     bnorm_floor (< 0: none) compared with sqrt(|b s|^2 / nlen); returns
 *     the device coefficient array (a, b_0 .. b_l) and the two skip flags.
 *   NVHip_LagStep: sums the K partials of pass p over nblk workgroups
 *     (NVHip_ReducePartials) in workgroup order and steps the state; last
 *     (see NVHip_LagNext) publishes a solve that goes on.
 *   NVHip_LagProjNorm: u = a w - sum_{j<nb} b_j U_j from the device
 *     coefficients, |u|^2 (u is not stored), then the last column's step.
 * One solve per stream at a time; returns 0 or -1. */
#define NVHIP_LAG_RUNNING (-1)
#define NVHIP_LAG_CONVERGED 0
#define NVHIP_LAG_MAXL 1
#define NVHIP_LAG_ZERO 2
#define NVHIP_LAG_BNORM 3
#define NVHIP_LAG_FALLBACK 4
#define NVHIP_LAG_QRFACT 5
typedef struct {
  int status;     /* NVHIP_LAG_* */
  int krydim;     /* Hessenberg columns formed */
  int iters;      /* Krylov iterations started */
  int passes;     /* passes run after pass 0 */
  double r_norm;  /* |u_0|, the preconditioned initial residual */
  double rho;     /* the final residual estimate */
  double nu[NVHIP_FUSED_MAX];                    /* |u_l| */
  double H[NVHIP_FUSED_MAX][NVHIP_FUSED_MAX];    /* Hessenberg, rows 0..maxl */
  double q[2 * NVHIP_FUSED_MAX];                 /* Givens rotations (c, s) */
} NVHipLagResult;
int NVHip_LagBegin(N_Vector v, int maxl, sunrealtype delta, sunrealtype bnorm_floor,
                   sunrealtype nlen, const double** coef_dev, const int** skip_dev,
                   const int** pskip_dev);
int NVHip_LagStep(N_Vector v, int K, int nblk, int p, int last);
int NVHip_LagProjNorm(N_Vector w, int nb, N_Vector* U);
int NVHip_LagWait(N_Vector v, NVHipLagResult* res);
/* a solve may be queued in parts: a step queued with last = 1 publishes
 * the state also when the solve goes on (status NVHIP_LAG_RUNNING); after
 * that wait NVHip_LagNext opens the next part (further passes, the
 * projection norm) with its own wait. */
int NVHip_LagNext(N_Vector v);
/* Queued behind the last part of a solve on b's stream: its solution
 * x = sum_k c_k V[k] / s (the coefficients c_k = y_k / |u_k| from the
 * solve's final step, which repeats the host exit's qr_sol and residual
 * test; rounded as NVHip_LinCombDiv) and then the first Newton update with
 * x (NVHip_NewtonUpdate with iter 0: ycor += x, y = zn0 + ycor, the sum of
 * (x w)^2; ycor must be zero on entry, as it is at a Newton iteration's
 * start, and is written without being read).  Both do nothing when the solve leaves no solution to use
 * (status other than converged or Krylov cap, a zero diagonal, or a cap
 * without residual reduction), and a part that ends with the solve still
 * running leaves none.  nvec: the basis vectors passed (>= the solve's
 * columns); x may be NULL when the caller does not read the solution (it
 * is then not written).  NVHip_LagUpdateWait waits for the update queued as *ticket:
 * dnorm = sqrt(sum / n), or -1 when the update did nothing.  No other
 * reduction of the stream may be waited for in between.  No reference
 * counterpart. */
int NVHip_LagUpdate(N_Vector b, int nvec, N_Vector* V, N_Vector s, N_Vector x, N_Vector ycor,
                    N_Vector zn0, N_Vector w, N_Vector y, unsigned* ticket);
int NVHip_LagUpdateWait(N_Vector ycor, unsigned ticket, sunrealtype* dnorm);

/* Reductions for kernels outside this file (no counterpart in the
 * reference): a kernel queued on v's stream leaves K partial sums per
 * workgroup at partials[k*nblk + b] (nblk*K <= *capacity, from
 * NVHip_ReducePartials); NVHip_ReduceFinish sums them over b < nblk in
 * workgroup order (deterministic), waits once and copies the K results to
 * out.  K <= NVHIP_FUSED_MAX.  Returns NULL / -1 for a non-HIP vector or
 * sizes beyond the capacity. */
double* NVHip_ReducePartials(N_Vector v, int* capacity);
int NVHip_ReduceFinish(N_Vector v, int K, int nblk, sunrealtype* out);
/* The same finish queued without waiting (skipped when *skip_dev is
 * nonzero on the device; skip_dev may be NULL); NVHip_ReduceWait waits for
 * it and copies the K results, or returns -1 when they never arrive (the
 * finish was skipped).  No other reduction of the stream may be waited for
 * in between. */
int NVHip_ReduceFinishAsync(N_Vector v, int K, int nblk, const int* skip_dev,
                            unsigned* ticket);
int NVHip_ReduceWait(N_Vector v, unsigned ticket, int K, sunrealtype* out);
/* A gate on v's stream for work queued before its scalar inputs are known
 * (no counterpart in the reference; it takes the kernel launch out of the
 * host round trip of an iterative solve).  NVHip_GateOpen queues a
 * one-thread kernel that waits until NVHip_GateRelease publishes up to 16
 * coefficients (copied to *coef_dev) or stop (*skip_dev = 1, which the
 * kernels queued behind it test to exit at once).  The wait is bounded
 * (1 s, then stop).  One gate per stream at a time; every opened gate must
 * be released, and nothing may synchronise the stream while one is open.
 * Requires NVHip_SetSyncMode 1 or 2.  Return 0 or -1. */
int NVHip_GateOpen(N_Vector v, const double** coef_dev, const int** skip_dev);
int NVHip_GateRelease(N_Vector v, int stop, const sunrealtype* coef, int n);
/* The projection norm of a classical Gram-Schmidt step with host
 * coefficients: *nn = |u|^2, u = a w - sum_{j<nb} b_j U_j, u written to
 * uout when uout is not NULL (uout may not alias w or any U_j). */
int NVHip_ProjNorm(int nb, sunrealtype a, const sunrealtype* b, N_Vector w, N_Vector* U,
                   N_Vector uout, sunrealtype* nn);

/* Fused-ops layer with the reference's entry-point names.
 * FusedNVec_Init installs the one-launch multi-dot and linear-combination
 * overrides on v and makes every clone of v inherit them;
 * FusedNVec_FreePool releases the reduction scratch pools. */
void FusedNVec_Init(N_Vector v);
void FusedNVec_FreePool(void);

#ifdef __cplusplus
}
#endif
#endif /* LLG_NVECTOR_HIP_H */
