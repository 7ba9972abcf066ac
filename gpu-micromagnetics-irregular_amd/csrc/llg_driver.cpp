// llg_driver.cpp -- the host driver of the reference's main(): wires the
// LLG plug-ins into the BDF integrator and runs the output loop
// (src/irregular/2d_fft/2d_i4/2d_fft.cu:545-700):
//   y/abstol vectors, FusedNVec_Init, CVodeInit(f), SVtolerances,
//   SPGMR(PREC_LEFT, KRYLOV_DIM) with classical GS iff neq < 500000,
//   SetJacTimes(JtvProduct), SetPreconditioner(PrecondSetup, PrecondSolve),
//   SetMaxOrd, then CVode(tout) every T1 with a timed region around the loop.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "llg/llg_cvode.h"
#include "llg/llg_driver.h"
#include "llg/llg_plugin.h"
#include "llg/nvector_hip.h"

namespace {

struct Run {
  llg_ctx* ctx = nullptr;
  N_Vector y = nullptr, abstol = nullptr;
  void* cv = nullptr;
  struct _SUNContext sctx{};
  llg_run_opts opts{};
  double t = 0.0;
};

}  // namespace

namespace {
// llgcv_lag_result and NVHipLagResult hold the same fields
int lag_wait(N_Vector v, llgcv_lag_result* out) {
  NVHipLagResult r;
  if (NVHip_LagWait(v, &r) != 0) return -1;
  static_assert(LLGCV_LAG_DIM == NVHIP_FUSED_MAX, "lagged SPGMR state dimensions");
  static_assert(LLGCV_LAG_CONVERGED == NVHIP_LAG_CONVERGED && LLGCV_LAG_MAXL == NVHIP_LAG_MAXL &&
                    LLGCV_LAG_ZERO == NVHIP_LAG_ZERO && LLGCV_LAG_BNORM == NVHIP_LAG_BNORM &&
                    LLGCV_LAG_FALLBACK == NVHIP_LAG_FALLBACK && LLGCV_LAG_QRFACT == NVHIP_LAG_QRFACT,
                "lagged SPGMR status codes");
  static_assert(LLGCV_LAG_RUNNING == NVHIP_LAG_RUNNING, "lagged SPGMR status codes");
  out->status = r.status;
  out->krydim = r.krydim;
  out->iters = r.iters;
  out->passes = r.passes;
  out->r_norm = r.r_norm;
  out->rho = r.rho;
  for (int i = 0; i < LLGCV_LAG_DIM; i++) out->nu[i] = r.nu[i];
  for (int i = 0; i < LLGCV_LAG_DIM; i++)
    for (int j = 0; j < LLGCV_LAG_DIM; j++) out->H[i][j] = r.H[i][j];
  for (int k = 0; k < 2 * LLGCV_LAG_DIM; k++) out->q[k] = r.q[k];
  return 0;
}
}  // namespace

extern "C" {

void LLG_DefaultRunOpts(llg_run_opts* o) {
  std::memset(o, 0, sizeof(*o));
  o->t0 = 0.0;
  o->t_out = 0.1;
  o->rtol = 1e-4;
  o->atol = 1e-4;
  o->maxl = 5;
  o->gstype = 0;        // auto: classical iff neq < 500000 (2d_i4/2d_fft.cu:645-650)
  o->max_order = 5;
  o->mxstep = 500;
  o->fused = 16151;
  o->use_jtv = 1;
  o->use_prec = 1;
  o->pretype = SUN_PREC_LEFT;
}

void* LLG_RunCreate(llg_ctx* ctx, const llg_run_opts* opts, const double* y0_host) {
  if (!ctx || !opts || !y0_host) return nullptr;
  Run* r = new Run();
  r->ctx = ctx;
  r->opts = *opts;
  int ncell = 0;
  LLG_GetCounts(ctx, &ncell, nullptr, nullptr);
  const sunindextype neq = 3 * (sunindextype)ncell;
  void* stream = LLG_GetStream(ctx);
  r->y = N_VNewWithStream_Hip(neq, stream, &r->sctx);
  r->abstol = N_VNewWithStream_Hip(neq, stream, &r->sctx);
  if (!r->y || !r->abstol) { LLG_RunDestroy(r); return nullptr; }
  if (opts->fused & 1) FusedNVec_Init(r->y);
  N_VSetFromHost_Hip(r->y, y0_host);
  N_VConst(opts->atol, r->abstol);
  r->cv = LLGCVodeCreate(LLGCV_BDF, &r->sctx);
  int gs = opts->gstype;
  if (gs == 0) gs = (neq < 500000) ? SUN_CLASSICAL_GS : SUN_MODIFIED_GS;
  if (!r->cv || LLGCVodeSetMaxOrd(r->cv, opts->max_order) != 0 ||
      LLGCVodeInit(r->cv, LLG_Rhs, opts->t0, r->y) != 0 ||
      LLGCVodeSetUserData(r->cv, ctx) != 0 ||
      LLGCVodeSVtolerances(r->cv, opts->rtol, r->abstol) != 0 ||
      LLGCVodeSetLinearSolverSPGMR(r->cv, opts->use_prec ? opts->pretype : SUN_PREC_NONE,
                                   opts->maxl, gs) != 0 ||
      LLGCVodeSetJacTimes(r->cv, nullptr, opts->use_jtv ? JtvProduct : nullptr) != 0 ||
      LLGCVodeSetMaxNumSteps(r->cv, opts->mxstep) != 0) {
    LLG_RunDestroy(r);
    return nullptr;
  }
  if (opts->use_prec) LLGCVodeSetPreconditioner(r->cv, PrecondSetup, PrecondSolve);
  if (opts->fused & 2) {
    llgcv_fused_ops fo{};
    fo.pascal = NVHip_Pascal;
    fo.ewt = NVHip_EwtSet;
    fo.scalediv = NVHip_ScaleDiv;
    fo.proddotmulti = NVHip_ProdDotMulti;
    fo.lincombdot = NVHip_LinCombDot;
    fo.gsproject = NVHip_GSProject;
    fo.lincombdiv = NVHip_LinCombDiv;
    fo.newtonupd = NVHip_NewtonUpdate;
    fo.predict_init = NVHip_PredictInit;
    fo.wrms_pair = NVHip_WrmsNormPair;
    if (opts->use_jtv) fo.atimes = LLG_Atimes;
    // the plug-in's jtimes and preconditioner callbacks never read fy
    if ((opts->fused & 256) && opts->use_jtv) {
      fo.rhs_residual = LLG_RhsResidual;
      fo.fy_free_jtimes = JtvProduct;
      fo.fy_free_psetup = opts->use_prec ? PrecondSetup : nullptr;
    }
    if (opts->fused & 512) {
      fo.predict_complete = NVHip_PredictComplete;
      fo.wrms_pair_comp = NVHip_WrmsNormPairComp;
      // bit 13: a completed step's copy of acor into zn[qmax] left to the
      // next predictor's pass as well
      if (opts->fused & 8192) fo.predict_complete_save = NVHip_PredictCompleteSave;
    }
    if ((opts->fused & 4) && opts->use_jtv && opts->use_prec && opts->pretype == SUN_PREC_LEFT) {
      fo.krylov = LLG_Krylov;
      fo.gsscaled = NVHip_GSProjectScaled;
      if (opts->fused & 8) fo.spgmr = LLG_SpgmrSolve;
      if (opts->fused & 16) {
        fo.arnoldi = LLG_Arnoldi;
        fo.projnorm = NVHip_ProjNorm;
      }
      if ((opts->fused & 80) == 80) fo.arnoldi_first = LLG_ArnoldiFirst;
      if ((opts->fused & 144) == 144) fo.newtonupd_comb = NVHip_NewtonUpdateComb;
      if ((opts->fused & 48) == 48) {
        fo.arnoldi_queue = LLG_ArnoldiQueue;
        fo.result_wait = NVHip_ReduceWait;
        fo.gate_open = NVHip_GateOpen;
        fo.gate_release = NVHip_GateRelease;
      }
      // bit 10: the lagged SPGMR queued whole, its scalar steps on the device
      if ((opts->fused & 1040) == 1040) {
        fo.lag_begin = NVHip_LagBegin;
        fo.arnoldi_lag = LLG_ArnoldiLag;
        fo.lag_projnorm = NVHip_LagProjNorm;
        fo.lag_wait = lag_wait;
        fo.lag_next = NVHip_LagNext;
        // bit 11: each solve's solution and first Newton update queued
        // behind it (one wait per Newton iteration instead of two)
        if (opts->fused & 2048) {
          fo.lag_update = NVHip_LagUpdate;
          fo.lag_update_wait = NVHip_LagUpdateWait;
          // bit 12: a step's order-change norms queued behind its first
          // Newton update (one wait less on such a step)
          if (opts->fused & 4096) {
            fo.pair_queue = NVHip_WrmsNormPairCompQueue;
            fo.pair_wait = NVHip_WrmsNormPairWait;
          }
        }
      }
    }
    LLGCVodeSetFusedOps(r->cv, &fo);
  }
  r->t = opts->t0;
  return r;
}

int LLG_RunAdvance(void* run, double t_end, llg_run_result* res) {
  Run* r = static_cast<Run*>(run);
  if (!r) return -1;
  hipStream_t s = (hipStream_t)LLG_GetStream(r->ctx);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipStreamSynchronize(s);
  const auto c0 = std::chrono::steady_clock::now();
  hipEventRecord(e0, s);
  const double t_start = r->t;
  double tout = r->t + r->opts.t_out;
  long nout = 0;
  int retval = 0;
  double tret = r->t;
  while (tout <= t_end + 1e-9 * std::fabs(t_end)) {
    retval = LLGCVode(r->cv, tout, r->y, &tret, LLGCV_NORMAL);
    if (retval < 0) break;
    nout++;
    r->t = tret;
    tout += r->opts.t_out;
  }
  hipEventRecord(e1, s);
  hipEventSynchronize(e1);
  const auto c1 = std::chrono::steady_clock::now();
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (res) {
    std::memset(res, 0, sizeof(*res));
    LLGCVodeGetStats(r->cv, &res->st);
    res->st.nsyncs = NVHip_NumSyncs();
    res->wall_s = std::chrono::duration<double>(c1 - c0).count();
    res->gpu_ms = ms;
    res->t_reached = r->t;
    res->sim_per_wall = res->wall_s > 0 ? (r->t - t_start) / res->wall_s : 0.0;
    res->retval = retval;
    res->nout = nout;
  }
  return retval < 0 ? retval : 0;
}

int LLG_RunGetState(void* run, double* y_host) {
  Run* r = static_cast<Run*>(run);
  if (!r || !y_host) return -1;
  return N_VGetToHost_Hip(r->y, y_host);
}

int LLG_RunStage(void* run, int which, const double* happ, double t_stage,
                 llg_run_result* res, double* m_avg) {
  Run* r = static_cast<Run*>(run);
  if (!r || !happ || !(t_stage > 0.0)) return -1;
  if (LLG_SetAppliedField(r->ctx, which, happ) != 0) return -1;
  if (LLGCVodeReInit(r->cv, r->t, r->y) != LLGCV_SUCCESS) return -1;
  const int rc = LLG_RunAdvance(run, r->t + t_stage, res);
  if (m_avg && LLG_MeanMagnetization(r->ctx, r->y, m_avg) != 0) return -1;
  return rc;
}

int LLG_RunRamp(void* run, int which, int nstages, const double* happ3, double t_stage,
                double* m_avg, llg_run_result* res) {
  Run* r = static_cast<Run*>(run);
  if (!r || !happ3 || nstages < 1) return -1;
  llg_run_result tot{};
  int rc = 0;
  for (int k = 0; k < nstages && rc == 0; k++) {
    llg_run_result one{};
    rc = LLG_RunStage(run, which, happ3 + 3 * k, t_stage, &one, m_avg ? m_avg + 3 * k : nullptr);
    llgcv_stats& a = tot.st;
    const llgcv_stats& b = one.st;
    a.nst += b.nst; a.nfe += b.nfe; a.nsetups += b.nsetups; a.nni += b.nni;
    a.ncfn += b.ncfn; a.netf += b.netf; a.nli += b.nli; a.nlcf += b.nlcf;
    a.njvevals += b.njvevals; a.npe += b.npe; a.nps += b.nps; a.nge += b.nge;
    a.nfeLS += b.nfeLS;
    a.qlast = b.qlast; a.qcur = b.qcur; a.hlast = b.hlast; a.hcur = b.hcur;
    a.tcur = b.tcur; a.h0u = b.h0u; a.nsyncs = b.nsyncs;
    tot.wall_s += one.wall_s;
    tot.gpu_ms += one.gpu_ms;
    tot.t_reached = one.t_reached;
    tot.nout += one.nout;
    tot.retval = one.retval;
  }
  const double span = (double)nstages * t_stage;
  tot.sim_per_wall = tot.wall_s > 0 ? span / tot.wall_s : 0.0;
  if (res) *res = tot;
  return rc;
}

void LLG_RunDestroy(void* run) {
  Run* r = static_cast<Run*>(run);
  if (!r) return;
  if (r->cv) LLGCVodeFree(&r->cv);
  N_VDestroy(r->y);
  N_VDestroy(r->abstol);
  delete r;
}

}  // extern "C"
