// kprof.h -- optional per-kernel timing with HIP events on the launch
// stream.  When a target kernel id is armed, every launch of that kernel
// (or one in `every`) is bracketed by an event pair from a pre-created
// pool; the total and the timed launch count are read back after the timed
// region.  Disarmed cost: one integer compare per launch.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace kprof {

enum Kernel {
  NONE = 0, RHS = 1, JTV = 2, BUILD_P = 3, APPLY_P = 4,
  DEMAG_ROW_FWD = 5, DEMAG_COL = 6, DEMAG_ROW_INV = 7, DEMAG_ALL = 8
};

struct State {
  int target = NONE;
  std::vector<hipEvent_t> ev;   // 2 per timed launch
  size_t used = 0;
  long dropped = 0;
  int every = 1;                // time one launch in `every` of the target
  long seen = 0;                // launches of the target so far
  bool open = false;            // begin() recorded, end() pending
};

State& state();

// whether this launch of the target is one of the timed sample (an event
// pair between two kernels costs the stream a few microseconds)
inline bool sampled(State& st) { return (st.seen++ % st.every) == 0; }

inline void begin(int id, hipStream_t s) {
  State& st = state();
  if (st.target != id) return;
  st.open = false;
  if (!sampled(st)) return;
  if (st.used + 2 > st.ev.size()) { st.dropped++; return; }
  hipEventRecord(st.ev[st.used], s);
  st.open = true;
}

// An event pair for one launch of kernel `id` made with
// hipExtLaunchKernelGGL, whose dispatch records both events itself (no
// marker packets between the kernels); false, events null, when the id is
// not armed, the launch is not in the sample or the pool is full.
inline bool ext_pair(int id, hipEvent_t* e0, hipEvent_t* e1) {
  *e0 = *e1 = nullptr;
  State& st = state();
  if (st.target != id) return false;
  if (!sampled(st)) return false;
  if (st.used + 2 > st.ev.size()) { st.dropped++; return false; }
  *e0 = st.ev[st.used];
  *e1 = st.ev[st.used + 1];
  st.used += 2;
  return true;
}

inline void end(int id, hipStream_t s) {
  State& st = state();
  if (st.target != id || !st.open) return;
  hipEventRecord(st.ev[st.used + 1], s);
  st.used += 2;
  st.open = false;
}

}  // namespace kprof
