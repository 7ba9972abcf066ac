// llg_ensemble.cpp -- concurrent independent instances (one context,
// stream, integrator and host thread each).  See llg_ensemble.h.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

#include "llg/llg_driver.h"
#include "llg/llg_ensemble.h"
#include "llg/llg_plugin.h"

namespace {

struct Ensemble {
  std::vector<hipStream_t> streams;
  std::vector<llg_ctx*> ctx;
  std::vector<void*> runs;
  std::vector<llg_run_result> last;
  std::vector<int> rc;
  int nthreads = 1;
};

void destroy(Ensemble* e) {
  if (!e) return;
  for (void* r : e->runs)
    if (r) LLG_RunDestroy(r);
  for (llg_ctx* c : e->ctx)
    if (c) LLG_Destroy(c);
  for (hipStream_t s : e->streams)
    if (s) hipStreamDestroy(s);
  delete e;
}

}  // namespace

extern "C" {

void* LLG_EnsembleCreate(int n, const llg_instance* inst, const llg_run_opts* opts,
                         int nthreads) {
  if (n < 1 || !inst || !opts) return nullptr;
  Ensemble* e = new Ensemble();
  e->nthreads = (nthreads <= 0 || nthreads > n) ? n : nthreads;
  e->streams.assign(n, nullptr);
  e->ctx.assign(n, nullptr);
  e->runs.assign(n, nullptr);
  e->last.assign(n, llg_run_result{});
  e->rc.assign(n, 0);
  for (int i = 0; i < n; i++) {
    if (!inst[i].y0 ||
        hipStreamCreateWithFlags(&e->streams[i], hipStreamNonBlocking) != hipSuccess) {
      destroy(e);
      return nullptr;
    }
    e->ctx[i] = LLG_Create(&inst[i].cfg, inst[i].active, inst[i].w, inst[i].cell_msk,
                           inst[i].cell_nsk, e->streams[i]);
    if (!e->ctx[i]) { destroy(e); return nullptr; }
    e->runs[i] = LLG_RunCreate(e->ctx[i], opts, inst[i].y0);
    if (!e->runs[i]) { destroy(e); return nullptr; }
    e->last[i].t_reached = opts->t0;
  }
  hipDeviceSynchronize();
  return e;
}

int LLG_EnsembleAdvance(void* ens, double t_end, llg_ensemble_result* res) {
  Ensemble* e = static_cast<Ensemble*>(ens);
  if (!e) return -1;
  const int n = (int)e->runs.size();
  std::vector<double> t_start(n);
  for (int i = 0; i < n; i++) t_start[i] = e->last[i].t_reached;
  const auto c0 = std::chrono::steady_clock::now();
  std::atomic<int> next{0};
  auto worker = [&]() {
    for (;;) {
      const int i = next.fetch_add(1);
      if (i >= n) return;
      e->rc[i] = LLG_RunAdvance(e->runs[i], t_end, &e->last[i]);
    }
  };
  std::vector<std::thread> th;
  th.reserve(e->nthreads);
  for (int k = 0; k < e->nthreads; k++) th.emplace_back(worker);
  for (auto& t : th) t.join();
  const auto c1 = std::chrono::steady_clock::now();
  if (res) {
    std::memset(res, 0, sizeof(*res));
    res->wall_s = std::chrono::duration<double>(c1 - c0).count();
    for (int i = 0; i < n; i++) {
      const llg_run_result& r = e->last[i];
      res->inst_time += r.t_reached - t_start[i];
      res->nst += r.st.nst;
      res->nfe += r.st.nfe;
      res->nli += r.st.nli;
      res->nni += r.st.nni;
      if (e->rc[i] != 0) res->failed++;
    }
  }
  for (int i = 0; i < n; i++)
    if (e->rc[i] != 0) return e->rc[i];
  return 0;
}

int LLG_EnsembleSize(void* ens) {
  Ensemble* e = static_cast<Ensemble*>(ens);
  return e ? (int)e->runs.size() : -1;
}

int LLG_EnsembleGetState(void* ens, int i, double* y_host) {
  Ensemble* e = static_cast<Ensemble*>(ens);
  if (!e || i < 0 || i >= (int)e->runs.size()) return -1;
  return LLG_RunGetState(e->runs[i], y_host);
}

int LLG_EnsembleGetResult(void* ens, int i, llg_run_result* r) {
  Ensemble* e = static_cast<Ensemble*>(ens);
  if (!e || !r || i < 0 || i >= (int)e->runs.size()) return -1;
  *r = e->last[i];
  return 0;
}

void LLG_EnsembleDestroy(void* ens) { destroy(static_cast<Ensemble*>(ens)); }

}  // extern "C"
