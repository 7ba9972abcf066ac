/*
 * llg_plugin.h -- the 2D LLG solver's CVODE plug-in surface on gfx950.
 *
 * Entry points and the reference interfaces they replace:
 *   LLG_Rhs        CVRhsFn f()            src/irregular/2d_fft/2d_i4/2d_fft.cu:328-366
 *                                         (compact: 2d_i6/2d_i6.cu:352-396,
 *                                          polycrystal: 2d_i3/2d_fft.cu:656-726)
 *   JtvProduct     CVLsJacTimesVecFn      2d_i4/jtv.h:26-27, jtv.cu:129-147
 *   PrecondSetup   CVLsPrecSetupFn        2d_i4/precond.h:31-33, precond.cu:324-358
 *                                         (two-phase jok: src/2d/2d_heter/precond.cu:265-314)
 *   PrecondSolve   CVLsPrecSolveFn        2d_i4/precond.h:35-38, precond.cu:361-385
 *   Precond_Create / Precond_Destroy      2d_i4/precond.h:28-29
 *   LLG_Create / LLG_Destroy              the UserData struct + its setup in main()
 *                                         (2d_i4/2d_fft.cu:219-233,472-543; 2d_i6/2d_i6.cu:214-231)
 *
 * `user_data` for every callback is an opaque llg_ctx* (instead of three
 * byte-mirrored structs).  The knobs of configs/sim_config.h become the
 * runtime struct llg_config.  Callbacks return 0 on success and -1 on a
 * launch failure (the reference's convention, 2d_i4/2d_fft.cu:359-364).
 * All kernels of a callback are launched on the HIP stream of the N_Vector
 * they write.
 */
#ifndef LLG_PLUGIN_H
#define LLG_PLUGIN_H

#include "sundials_compat.h"
#include "demag_fft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* execution model (configs/sim_config.h:107-108) */
#define LLG_EXEC_YMSK     0
#define LLG_EXEC_COMPACT  1
/* anisotropy form (configs/sim_config.h:279-280) */
#define LLG_ANISO_LINEAR  0   /* h_a += msk_a (chk (m.msk) + cha)      (i1, i3) */
#define LLG_ANISO_CUBIC   1   /* h_a += msk_a chk m_a (m_a^2 - 1)      (i2, i4, i5, i6) */
/* normalisation of y at the top of f (A4) */
#define LLG_NORM_NONE         0   /* i4, i5, i6 */
#define LLG_NORM_EXACT        1   /* i1: m/|m| when |m| > 1e-30 */
#define LLG_NORM_REGULARIZED  2   /* i2, i3: m/(|m| + 1e-3) */
/* lattice boundary */
#define LLG_BC_PERIODIC   0   /* all reference variants */
#define LLG_BC_NEUMANN    1   /* mirror ghost cell (the 128x128 plain workload) */
/* demag */
#define LLG_DEMAG_OFF       0
#define LLG_DEMAG_PERIODIC  1   /* reference operator */
#define LLG_DEMAG_PADDED    2   /* open boundary, zero-padded 2nx x 2ny */
/* output gate of a callback */
#define LLG_GATE_NONE    0
#define LLG_GATE_MASK    1   /* x ymsk (0/1 per cell) */
#define LLG_GATE_WINDOW  2   /* x w (soft window, i3) */
/* preconditioner setup protocol */
#define LLG_PC_ALWAYS   0   /* rebuild every setup, *jcur = TRUE (i4/i6) */
#define LLG_PC_JOK      1   /* build J only when !jok, P^-1 every setup (2d_heter) */

/* One set of field terms.  RHS, Jv and the preconditioner each carry
 * their own set because the reference's three mirrors differ
 * (e.g. i4 pc_happ1 = -1.0 vs happ1 = -0.2; i4 Jv has no Zeeman). */
typedef struct llg_terms {
  double che;        /* exchange coupling */
  double chk;        /* anisotropy */
  double cha;        /* linear-anisotropy offset */
  double chb;        /* DMI-like coupling to the x neighbours */
  double msk[3];     /* uniform anisotropy axis */
  double nsk[3];     /* uniform DMI direction */
  double happ[3];    /* Zeeman field */
  int    aniso;      /* LLG_ANISO_* */
  int    cell_axes;  /* 1: per-cell axes (LLG_Create cell_msk/cell_nsk) override msk/nsk */
  int    window;     /* 1: neighbour m enters as w*m (i3 RHS) */
  int    gate;       /* LLG_GATE_* */
} llg_terms;

typedef struct llg_config {
  int    ng, ny;           /* cell grid (reference ng = NX_VAL/3, ny) */
  int    exec;             /* LLG_EXEC_* */
  int    boundary;         /* LLG_BC_* */
  int    exch_self;        /* 1: exchange = che * (sum_nb m_nb - n_nb m) (Laplacian form) */
  int    normalize;        /* LLG_NORM_* */
  int    demag;            /* LLG_DEMAG_* */
  int    demag_fp32;       /* 1: FP32 spectral demag */
  int    pc_mode;          /* LLG_PC_* */
  int    pad0;
  double demag_strength;
  double demag_thick;
  double alpha;            /* damping */
  double chg;              /* gyromagnetic prefactor */
  llg_terms rhs, jv, pc;
} llg_config;

typedef struct llg_ctx llg_ctx;

/* Build a context (= the reference's UserData + its device buffers).
 * active   : ng*ny bytes, nonzero = magnetic cell; NULL = all active
 * w        : ng*ny soft window (NULL = none)
 * cell_msk : 3*ng*ny SoA per-cell anisotropy axes (NULL = uniform)
 * cell_nsk : 3*ng*ny SoA per-cell DMI directions  (NULL = uniform)
 * stream   : hipStream_t for the context's own work (NULL = default)     */
llg_ctx* LLG_Create(const llg_config* cfg, const unsigned char* active,
                    const double* w, const double* cell_msk,
                    const double* cell_nsk, void* stream);
void     LLG_Destroy(llg_ctx* ctx);
const char* LLG_LastError(void);

/* accessors */
int      LLG_GetCounts(llg_ctx* ctx, int* ncell, int* n_active, int* n_inactive);
int      LLG_GetMask(llg_ctx* ctx, double* h_ymsk3);            /* 3*ncell, 1.0/0.0 */
int      LLG_GetActiveIds(llg_ctx* ctx, int* h_ids, int* h_inactive_ids);
int      LLG_GetHdemag(llg_ctx* ctx, double* h_hd3);            /* last f()'s field */
int      LLG_GetPinv(llg_ctx* ctx, double* h_p9);               /* 9*ncell SoA */
DemagData* LLG_GetDemag(llg_ctx* ctx);
void*    LLG_GetStream(llg_ctx* ctx);
int      LLG_GetGrid(llg_ctx* ctx, int* ng, int* ny);
/* 2D workgroup shape of the stencil kernels (RHS, Jv, P): bx*by a multiple
 * of 64, at most 256.  Results do not depend on it. */
int      LLG_SetStencilBlock(llg_ctx* ctx, int bx, int by);
/* Autotuner: times the Jv stencil over workgroup shapes and the demag
 * passes over FFT kernel variants, workgroup sizes and LDS paddings
 * (Demag_Autotune), keeps the fastest.  choice[8] = {bx, by, demag row
 * block, col block, row pad, col pad, row variant, col variant}.  No
 * reference counterpart (the reference fixes 256-thread 1D launches at
 * compile time). */
int      LLG_Autotune(llg_ctx* ctx, int reps, int* choice);
/* device copy of the active flags (one byte per cell) */
const unsigned char* LLG_GetActiveDevice(llg_ctx* ctx);
/* per-callback launch counters (diagnostics) */
int      LLG_GetCallCounts(llg_ctx* ctx, long* nf, long* njv, long* npsetup, long* npsolve);

/* Per-kernel timing with HIP events on the launch stream (diagnostics; no
 * reference counterpart).  kernel_id: 1 RHS, 2 Jv, 3 block-Jacobi setup,
 * 4 block-Jacobi solve, 5 demag row forward, 6 demag column, 7 demag row
 * inverse, 8 whole demag.  capacity = launches to record. */
int LLG_ProfileStart(int kernel_id, int capacity);
/* time one launch in `every` of the armed kernel (an event pair between two
 * kernels delays the next dispatch by a few microseconds); after
 * LLG_ProfileStart, which resets it to 1 */
int LLG_ProfileSetEvery(int every);
int LLG_ProfileStop(double* total_ms, long* launches, long* dropped);
/* HIP event helpers for timing from a host binding (diagnostics) */
void*  LLG_EventCreate(void);
int    LLG_EventRecord(void* ev, void* stream);
double LLG_EventElapsedMs(void* e0, void* e1);
void   LLG_EventDestroy(void* ev);
int    LLG_DeviceSynchronize(void);

/* Term sets of a context (bit mask for LLG_SetAppliedField) */
#define LLG_TERMS_RHS 1
#define LLG_TERMS_JV  2
#define LLG_TERMS_PC  4
/* Set the applied (Zeeman) field of the term sets in `which` (a mask of
 * LLG_TERMS_*), e.g. one stage of a quasi-static hysteresis sweep; it takes
 * effect at the next callback.  The reference's field is a compile-time
 * constant in each of its three mirrors (2d_i4/2d_fft.cu:178-180 happ1..3,
 * precond.cu:68-70 pc_happ1..3; its Jv has none, jtv.cu:36-39), so a field
 * change there means a rebuild.  Returns 0, or -1 on a NULL argument or a
 * mask outside 1..7.  LLG_GetAppliedField reads one set (which = one bit). */
int LLG_SetAppliedField(llg_ctx* ctx, int which, const double* happ);
int LLG_GetAppliedField(llg_ctx* ctx, int which, double* happ);
/* Mean magnetisation over the active cells of y (3 values), reduced in a
 * fixed order on the device (the same bits on every call); one host wait.
 * No reference counterpart (the reference writes whole frames,
 * 2d_i4/2d_fft.cu:393-411, and averages them in its plotting scripts). */
int LLG_MeanMagnetization(llg_ctx* ctx, N_Vector y, double* m);

/* CVODE plug-ins (user_data = llg_ctx*) */
int LLG_Rhs(sunrealtype t, N_Vector y, N_Vector ydot, void* user_data);
/* The right-hand side and the Newton residual built from it in one RHS
 * pass: res = c[0] zn1 + c[1] ycor + c[2] f(t, y), equal bit for bit to
 * LLG_Rhs into a vector f followed by N_VLinearCombination(3, c,
 * {zn1, ycor, f}, res) (CVODE's cvNlsResidual); f itself is not written.
 * res must not alias y, zn1 or ycor.  Returns 0, -1 on error, or 2 without
 * doing anything when the grid runs in compact mode (then call LLG_Rhs and
 * combine).  No reference counterpart: installed as the integrator's fused
 * Newton residual. */
int LLG_RhsResidual(sunrealtype t, N_Vector y, const sunrealtype* c, N_Vector zn1,
                    N_Vector ycor, N_Vector res, void* user_data);
int JtvProduct(N_Vector v, N_Vector Jv, sunrealtype t, N_Vector y, N_Vector fy,
               void* user_data, N_Vector tmp);
/* The SPGMR operator of the Newton system in one launch:
 * z = v - gamma * J(y) v, equal bit for bit to JtvProduct followed by
 * N_VLinearSum(1, v, -gamma, Jv, z) (CVODE's cvLsATimes).  z must not
 * alias v.  No reference counterpart: installed as the integrator's fused
 * atimes by the driver (llgcv_fused_ops.atimes). */
int LLG_Atimes(N_Vector v, N_Vector z, sunrealtype gamma, sunrealtype t, N_Vector y,
               N_Vector fy, void* user_data);
/* One SPGMR iteration's operator chain with left block-Jacobi
 * preconditioning, in one launch: with u = (c vl) / s (formed where read;
 * vl is not written), vk = s .* P^{-1}(u - gamma J(y) u).  Equal bit for
 * bit to N_VScale(c, vl, vl), N_VDiv(vl, s, u), LLG_Atimes(u, z),
 * PrecondSolve(z, t) and N_VProd(s, t, vk) (the CVLS/SPGMR sequence,
 * apart from vl staying unscaled).  Returns -1 without launching when a
 * learned correction is active or vk aliases vl, s or y.  No reference
 * counterpart: installed as llgcv_fused_ops.krylov by the driver. */
int LLG_Krylov(sunrealtype c, N_Vector vl, N_Vector s, N_Vector vk, sunrealtype gamma,
               sunrealtype t, N_Vector y, N_Vector fy, void* user_data);
/* A whole left-preconditioned SPGMR solve (classical GS) queued on the
 * device with one host wait (NVHip_SpgmrSolve with the LLG_Krylov chain as
 * its operator): t0 = V[0] holds P^{-1} b on entry, x receives the
 * solution, *status one of NVHIP_SPGMR_*.  Same bits as the host-driven
 * solve.  Returns -1 without queuing anything when a learned correction is
 * active or no block-Jacobi data exists.  No reference counterpart:
 * installed as llgcv_fused_ops.spgmr by the driver. */
int LLG_SpgmrSolve(N_Vector t0, N_Vector* V, int maxl, N_Vector s, N_Vector x,
                   sunrealtype delta, sunrealtype gamma, sunrealtype t, N_Vector y, N_Vector fy,
                   void* user_data, int* status, int* iters, sunrealtype* r_norm,
                   sunrealtype* resnorm);
/* One iteration of left-preconditioned SPGMR with the normalisation of the
 * basis vector left to the caller (A = s .* P^{-1}(I - gamma J(y)) (. / s)
 * is linear, so A(u/|u|) = (A u)/|u|), in one pass on the dense grid:
 *   u = a win - sum_{j<l} b_j U[j]   (l > 0; the classical Gram-Schmidt
 *                                     projection of the previous step),
 *     = s .* P^{-1} win              (l == 0, pre != 0: the first
 *                                     preconditioner solve and scaling),
 *     = U[0]                         (l == 0, pre == 0);
 *   u written to U[l] (l > 0 or pre), wout = A u, and one host wait for
 *   r[0] = |u|^2, r[1+j] = <wout, U[j]> (j < l), r[1+l] = <wout, u>,
 *   r[2+l] = |wout|^2, and with pre r[3] = sum_i (win_i s_i)^2 (the square
 *   of |win|_WRMS(s) times the length, the right-hand-side norm CVLS checks).
 * l < 8.  U[l] and wout may not alias win, s, y or each other.  Returns -1
 * without launching on compact execution, with a learned correction
 * active, or on aliasing.  No reference counterpart: installed as
 * llgcv_fused_ops.arnoldi by the driver (SUNDIALS' SPGMR waits for the
 * multi-dot and for the norm separately). */
int LLG_Arnoldi(int l, int pre, sunrealtype a, const sunrealtype* b, N_Vector win, N_Vector* U,
                N_Vector wout, N_Vector s, sunrealtype gamma, sunrealtype t, N_Vector y,
                N_Vector fy, void* user_data, sunrealtype* r);
/* LLG_Arnoldi queued without waiting: *ticket for NVHip_ReduceWait(wout,
 * ticket, l + 3 + (pre != 0), r).  With coef_dev, a and b are read from
 * device memory (coef_dev[0] = a, coef_dev[1 + j] = b_j) when the kernel
 * runs; with skip_dev, the pass and its finish exit at once when
 * *skip_dev is nonzero (the coefficients and the stop flag of an
 * NVHip_GateOpen gate).  No reference counterpart. */
int LLG_ArnoldiQueue(int l, int pre, sunrealtype a, const sunrealtype* b, const double* coef_dev,
                     const int* skip_dev, N_Vector win, N_Vector* U, N_Vector wout, N_Vector s,
                     sunrealtype gamma, sunrealtype t, N_Vector y, N_Vector fy, void* user_data,
                     unsigned* ticket);
/* LLG_Arnoldi queued for the device-driven SPGMR step (NVHip_LagBegin):
 * pass l (l = 0 with pre: the preconditioned right-hand side win = b;
 * l > 0: u_l = a win - sum_{j<l} b_j U_j with a, b read from coef_dev,
 * the pass and its step exiting at once when *skip_dev is nonzero), its
 * sums summed by NVHip_LagStep(wout, ..., l, last) on the device.  Nothing
 * is waited for.  No reference counterpart. */
int LLG_ArnoldiLag(int l, int pre, int last, const double* coef_dev, const int* skip_dev,
                   N_Vector win, N_Vector* U, N_Vector wout, N_Vector s, sunrealtype gamma,
                   sunrealtype t, N_Vector y, N_Vector fy, void* user_data);
/* The first pass (LLG_Arnoldi with l = 0, pre) with the right-hand side
 * formed where it is read: b = sum_{k<nb} c_k X_k (nb <= 3, rounded as
 * N_VLinearCombination rounds it; the Newton residual), so b is never
 * written.  r as for LLG_Arnoldi (4 sums).  No reference counterpart. */
int LLG_ArnoldiFirst(int nb, const sunrealtype* c, N_Vector* X, N_Vector* U, N_Vector wout,
                     N_Vector s, sunrealtype gamma, sunrealtype t, N_Vector y, N_Vector fy,
                     void* user_data, sunrealtype* r);
int PrecondSetup(sunrealtype t, N_Vector y, N_Vector fy, sunbooleantype jok,
                 sunbooleantype* jcurPtr, sunrealtype gamma, void* user_data);
int PrecondSolve(sunrealtype t, N_Vector y, N_Vector fy, N_Vector r, N_Vector z,
                 sunrealtype gamma, sunrealtype delta, int lr, void* user_data);

/* Block-Jacobi storage (reference-compatible handle) */
typedef struct PrecondData PrecondData;
PrecondData* Precond_Create(int ng, int ny, int ncell);
void         Precond_Destroy(PrecondData* pd);

#ifdef __cplusplus
}
#endif
#endif /* LLG_PLUGIN_H */
