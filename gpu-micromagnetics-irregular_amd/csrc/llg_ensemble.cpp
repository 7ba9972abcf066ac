// llg_ensemble.cpp -- concurrent independent instances (one context,
// stream, integrator and host thread each), optionally with the f() calls
// of same-shape instances gathered into one batched evaluation.  See
// llg_ensemble.h.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "llg/llg_batch.h"
#include "llg/llg_driver.h"
#include "llg/llg_ensemble.h"
#include "llg/llg_plugin.h"
#include "llg_hook.h"

namespace {

// mode 2 batches the groups of at most this many cells per instance
constexpr long kBatchSmallCells = 1L << 17;

struct Ensemble;

// Instances of one grid shape whose f() calls meet: each member's call
// records an event on its stream and waits; the last live member to arrive
// launches one LLG_RhsBatch for all of them on the group's stream (behind
// every member's event) and records `done`, which each member's stream then
// waits for.  A member that finishes its advance leaves the group.  Every
// member's calls depend only on its own stream and on batches already
// launched, so the group never waits on itself.
struct Group {
  std::vector<int> members;        // ensemble indices; slot k = members[k]
  llg_batch* batch = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  std::mutex mu;
  std::condition_variable cv;
  int live = 0, arrived = 0, last_rc = 0;
  long gen = 0;
  std::vector<double*> y, out;
  std::vector<llg_res_in> res;
  std::vector<const llg_res_in*> resp;
  long launches = 0, calls = 0;
};

struct Member {
  Ensemble* e = nullptr;
  Group* g = nullptr;
  int index = 0, slot = 0;
  hipEvent_t ready = nullptr;
};

struct Ensemble {
  std::vector<hipStream_t> streams;
  std::vector<llg_ctx*> ctx;
  std::vector<void*> runs;
  std::vector<llg_run_result> last;
  std::vector<int> rc;
  int nthreads = 1;
  std::vector<std::unique_ptr<Group>> groups;
  std::vector<Member> members;      // per instance (batched mode)
};

// with g.mu held: evaluate f for every member that arrived
void launch_locked(Ensemble* e, Group& g) {
  const int n = (int)g.members.size();
  for (int k = 0; k < n; k++)
    if (g.y[k]) hipStreamWaitEvent(g.stream, e->members[g.members[k]].ready, 0);
  int rc = LLG_RhsBatch(g.batch, g.y.data(), g.out.data(), g.resp.data(), g.stream);
  if (rc == 0 && hipEventRecord(g.done, g.stream) != hipSuccess) rc = -1;
  g.launches++;
  g.calls += g.arrived;
  for (int k = 0; k < n; k++) {
    g.y[k] = nullptr;
    g.out[k] = nullptr;
    g.resp[k] = nullptr;
  }
  g.arrived = 0;
  g.last_rc = rc;
  g.gen++;
  g.cv.notify_all();
}

int group_rhs(void* data, llg_ctx*, double* y, double* ydot, const llg_res_in* res) {
  Member* m = static_cast<Member*>(data);
  Ensemble* e = m->e;
  Group& g = *m->g;
  hipStream_t ms = e->streams[m->index];
  if (hipEventRecord(m->ready, ms) != hipSuccess) return -1;
  int rc;
  {
    std::unique_lock<std::mutex> lk(g.mu);
    g.y[m->slot] = y;
    g.out[m->slot] = ydot;
    if (res) {
      g.res[m->slot] = *res;
      g.resp[m->slot] = &g.res[m->slot];
    }
    g.arrived++;
    const long my = g.gen;
    if (g.arrived == g.live) launch_locked(e, g);
    else g.cv.wait(lk, [&] { return g.gen != my; });
    rc = g.last_rc;
  }
  // this member's later work reads the batch's output; `done` cannot be
  // recorded again before this member arrives once more
  if (hipStreamWaitEvent(ms, g.done, 0) != hipSuccess) return -1;
  return rc;
}

void group_leave(Member* m) {
  Group& g = *m->g;
  std::lock_guard<std::mutex> lk(g.mu);
  g.live--;
  if (g.arrived > 0 && g.arrived == g.live) launch_locked(m->e, g);
}

void clear_groups(Ensemble* e) {
  for (size_t i = 0; i < e->ctx.size(); i++)
    if (e->ctx[i]) LLG_SetRhsHook(e->ctx[i], nullptr, nullptr);
  for (Member& m : e->members)
    if (m.ready) hipEventDestroy(m.ready);
  e->members.clear();
  for (auto& g : e->groups) {
    if (g->batch) LLG_BatchDestroy(g->batch);
    if (g->done) hipEventDestroy(g->done);
    if (g->stream) hipStreamDestroy(g->stream);
  }
  e->groups.clear();
}

void destroy(Ensemble* e) {
  if (!e) return;
  clear_groups(e);
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
  const bool batched = !e->groups.empty();
  for (auto& g : e->groups) {
    g->live = (int)g->members.size();
    g->arrived = 0;
  }
  std::atomic<int> next{0};
  auto worker = [&]() {
    for (;;) {
      const int i = next.fetch_add(1);
      if (i >= n) return;
      e->rc[i] = LLG_RunAdvance(e->runs[i], t_end, &e->last[i]);
      if (batched && e->members[i].g) group_leave(&e->members[i]);
    }
  };
  // batched: every instance needs a thread of its own (a group waits for
  // all its live members)
  const int nth = batched ? n : e->nthreads;
  std::vector<std::thread> th;
  th.reserve(nth);
  for (int k = 0; k < nth; k++) th.emplace_back(worker);
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

int LLG_EnsembleSetBatched(void* ens, int mode) {
  Ensemble* e = static_cast<Ensemble*>(ens);
  if (!e || mode < 0 || mode > 2) return -1;
  hipDeviceSynchronize();
  clear_groups(e);
  if (mode == 0) return 0;
  const int n = (int)e->ctx.size();
  std::map<std::pair<int, int>, std::vector<int>> by_shape;
  for (int i = 0; i < n; i++) {
    int ng = 0, ny = 0;
    LLG_GetGrid(e->ctx[i], &ng, &ny);
    by_shape[{ng, ny}].push_back(i);
  }
  e->members.assign(n, Member{});
  for (auto& kv : by_shape) {
    // mode 2: only the small grids, whose instances are bound by kernel
    // dispatch rather than by memory traffic
    if (mode == 2 && (long)kv.first.first * kv.first.second > kBatchSmallCells) continue;
    auto g = std::make_unique<Group>();
    g->members = kv.second;
    const int m = (int)g->members.size();
    std::vector<llg_ctx*> cs(m);
    for (int k = 0; k < m; k++) cs[k] = e->ctx[g->members[k]];
    g->batch = LLG_BatchCreate(m, cs.data());
    if (!g->batch ||
        hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&g->done, hipEventDisableTiming) != hipSuccess) {
      e->groups.push_back(std::move(g));
      clear_groups(e);
      return -1;
    }
    g->y.assign(m, nullptr);
    g->out.assign(m, nullptr);
    g->res.assign(m, llg_res_in{});
    g->resp.assign(m, nullptr);
    for (int k = 0; k < m; k++) {
      Member& mb = e->members[g->members[k]];
      mb.e = e;
      mb.g = g.get();
      mb.index = g->members[k];
      mb.slot = k;
    }
    e->groups.push_back(std::move(g));
  }
  for (int i = 0; i < n; i++) {
    if (!e->members[i].g) continue;   // evaluates f on its own
    if (hipEventCreateWithFlags(&e->members[i].ready, hipEventDisableTiming) != hipSuccess) {
      clear_groups(e);
      return -1;
    }
    LLG_SetRhsHook(e->ctx[i], group_rhs, &e->members[i]);
  }
  return (int)e->groups.size();
}

int LLG_EnsembleBatchStats(void* ens, long* launches, long* calls) {
  Ensemble* e = static_cast<Ensemble*>(ens);
  if (!e) return -1;
  long l = 0, c = 0;
  for (auto& g : e->groups) {
    l += g->launches;
    c += g->calls;
  }
  if (launches) *launches = l;
  if (calls) *calls = c;
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
