// Per-phase GPU timing with HIP events recorded on the caller's stream around
// each phase's launches (off unless rts_profile_enable(1)).
#include <mutex>
#include <vector>

#include "rts_internal.h"

namespace rts {
namespace {

struct Pending {
  int phase;
  hipEvent_t begin, end;
};

struct Profiler {
  std::mutex mu;
  bool on = false;
  std::vector<Pending> pending;
  std::vector<hipEvent_t> spare;
  double ms[RTS_PHASE_COUNT] = {};
  uint64_t count[RTS_PHASE_COUNT] = {};
};

Profiler& prof() {
  static Profiler* p = new Profiler();  // never destroyed: events outlive static teardown
  return *p;
}

hipEvent_t take_event(Profiler& p) {
  if (!p.spare.empty()) {
    hipEvent_t e = p.spare.back();
    p.spare.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

thread_local int g_phases_off = 0;

}  // namespace

PhasesOff::PhasesOff() { g_phases_off++; }
PhasesOff::~PhasesOff() { g_phases_off--; }

PhaseToken phase_begin(int phase, hipStream_t stream) {
  PhaseToken t{-1, nullptr};
  if (g_phases_off) return t;
  Profiler& p = prof();
  std::lock_guard<std::mutex> lk(p.mu);
  if (!p.on || phase < 0 || phase >= RTS_PHASE_COUNT) return t;
  hipEvent_t e = take_event(p);
  if (!e || hipEventRecord(e, stream) != hipSuccess) return t;
  t.phase = phase;
  t.begin = e;
  return t;
}

void phase_end(const PhaseToken& t, hipStream_t stream) {
  if (t.phase < 0) return;
  Profiler& p = prof();
  std::lock_guard<std::mutex> lk(p.mu);
  hipEvent_t e = take_event(p);
  if (!e || hipEventRecord(e, stream) != hipSuccess) {
    p.spare.push_back(t.begin);
    return;
  }
  p.pending.push_back(Pending{t.phase, t.begin, e});
}

}  // namespace rts

extern "C" {

void rts_profile_enable(int on) {
  rts::Profiler& p = rts::prof();
  std::lock_guard<std::mutex> lk(p.mu);
  p.on = on != 0;
}

int rts_profile_read(rts_profile* out, int reset) {
  rts::Profiler& p = rts::prof();
  std::lock_guard<std::mutex> lk(p.mu);
  for (auto& e : p.pending) {
    float ms = 0.f;
    if (hipEventSynchronize(e.end) == hipSuccess &&
        hipEventElapsedTime(&ms, e.begin, e.end) == hipSuccess) {
      p.ms[e.phase] += ms;
      p.count[e.phase] += 1;
    }
    p.spare.push_back(e.begin);
    p.spare.push_back(e.end);
  }
  p.pending.clear();
  if (out) {
    for (int i = 0; i < RTS_PHASE_COUNT; i++) {
      out->ms[i] = p.ms[i];
      out->count[i] = p.count[i];
    }
  }
  if (reset) {
    for (int i = 0; i < RTS_PHASE_COUNT; i++) {
      p.ms[i] = 0;
      p.count[i] = 0;
    }
  }
  return RTS_OK;
}

}  // extern "C"
