// kprof.h -- optional per-kernel timing with HIP events on the launch
// stream.  When a target kernel id is armed, every launch of that kernel
// is bracketed by an event pair from a pre-created pool; the total and the
// launch count are read back after the timed region.  Disarmed cost: one
// integer compare per launch.
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
  std::vector<hipEvent_t> ev;   // 2 per launch
  size_t used = 0;
  long dropped = 0;
};

State& state();

inline void begin(int id, hipStream_t s) {
  State& st = state();
  if (st.target != id) return;
  if (st.used + 2 > st.ev.size()) { st.dropped++; return; }
  hipEventRecord(st.ev[st.used], s);
}

// An event pair for one launch of kernel `id` made with
// hipExtLaunchKernelGGL, whose dispatch records both events itself (no
// marker packets between the kernels); false, events null, when the id is
// not armed or the pool is full.
inline bool ext_pair(int id, hipEvent_t* e0, hipEvent_t* e1) {
  *e0 = *e1 = nullptr;
  State& st = state();
  if (st.target != id) return false;
  if (st.used + 2 > st.ev.size()) { st.dropped++; return false; }
  *e0 = st.ev[st.used];
  *e1 = st.ev[st.used + 1];
  st.used += 2;
  return true;
}

inline void end(int id, hipStream_t s) {
  State& st = state();
  if (st.target != id) return;
  if (st.used + 2 > st.ev.size()) return;
  hipEventRecord(st.ev[st.used + 1], s);
  st.used += 2;
}

}  // namespace kprof
