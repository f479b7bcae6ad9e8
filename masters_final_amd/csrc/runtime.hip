// Error reporting and small host-side runtime pieces shared by every entry
// point of libeac_hip.so.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "common.hpp"

namespace eac {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

// One pinned block per host thread; grown on demand, never freed before
// process exit (pinned memory is allocated, never registered).
struct Pinned {
  void* p = nullptr;
  size_t bytes = 0;
};
static thread_local Pinned g_pinned;

void* pinned_scratch(size_t bytes) {
  if (g_pinned.bytes >= bytes) return g_pinned.p;
  size_t want = bytes < 4096 ? 4096 : bytes;
  void* p = nullptr;
  if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) return nullptr;
  if (g_pinned.p) (void)hipHostFree(g_pinned.p);
  g_pinned.p = p;
  g_pinned.bytes = want;
  return p;
}

// ---------------------------------------------------------------- profiler
struct ProfSlot {
  double ms = 0.0;
  int64_t launches = 0;
  int64_t bytes = 0;
};
// Profiler state is per host thread: a thread's records, totals and event
// pool are its own, so calls from several host threads never share them.
static thread_local bool g_prof = false;
static thread_local bool g_prof_suspended = false;  // set while a hipGraph is captured
static thread_local ProfSlot g_slots[EAC_PROF_SLOTS];
static thread_local std::vector<ProfRec> g_recs;
static thread_local std::vector<hipEvent_t> g_pool;

bool prof_on() { return g_prof && !g_prof_suspended; }

void prof_suspend(bool on) { g_prof_suspended = on; }

static hipEvent_t ev_get() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

int prof_begin(int slot, int64_t bytes, hipStream_t s) {
  if (!prof_on() || slot < 0 || slot >= EAC_PROF_SLOTS) return -1;
  ProfRec r{slot, ev_get(), ev_get(), bytes};
  if (!r.a || !r.b) return -1;
  (void)hipEventRecord(r.a, s);
  g_recs.push_back(r);
  return (int)g_recs.size() - 1;
}

void prof_end(int rec, hipStream_t s) {
  if (rec < 0 || rec >= (int)g_recs.size()) return;
  (void)hipEventRecord(g_recs[rec].b, s);
}

void prof_set_bytes(int rec, int64_t bytes) {
  if (rec >= 0 && rec < (int)g_recs.size()) g_recs[rec].bytes = bytes;
}

void prof_flush() {
  for (auto& r : g_recs) {
    float ms = 0.f;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      g_slots[r.slot].ms += ms;
      g_slots[r.slot].launches += 1;
      g_slots[r.slot].bytes += r.bytes;
    }
    g_pool.push_back(r.a);
    g_pool.push_back(r.b);
  }
  g_recs.clear();
}

}  // namespace eac

extern "C" int eac_profile_enable(int32_t on) {
  eac::prof_flush();
  eac::g_prof = on != 0;
  for (auto& s : eac::g_slots) s = eac::ProfSlot();
  return EAC_OK;
}

extern "C" int eac_profile_read(int32_t slot, double* total_ms, int64_t* launches,
                                int64_t* alg_bytes) {
  if (slot < 0 || slot >= EAC_PROF_SLOTS) return eac::fail(EAC_ERR_ARG, "profile: bad slot");
  eac::prof_flush();
  if (total_ms) *total_ms = eac::g_slots[slot].ms;
  if (launches) *launches = eac::g_slots[slot].launches;
  if (alg_bytes) *alg_bytes = eac::g_slots[slot].bytes;
  return EAC_OK;
}

extern "C" const char* eac_last_error(void) { return eac::g_err; }

extern "C" int eac_abi_version(void) { return 1; }

// ------------------------------------------------------------ tuning knobs
namespace eac {
// Per host thread: a setting applies to the calls its own thread makes.
static thread_local int g_tune[EAC_TUNE_COUNT] = {0};
int tune_value(int knob) { return (knob >= 0 && knob < EAC_TUNE_COUNT) ? g_tune[knob] : 0; }
}  // namespace eac

extern "C" int eac_tune_set(int32_t knob, int32_t value) {
  if (knob < 0 || knob >= EAC_TUNE_COUNT) return eac::fail(EAC_ERR_ARG, "tune: bad knob %d", knob);
  if (value < -1 || value > (1 << 20)) return eac::fail(EAC_ERR_ARG, "tune: bad value %d", value);
  eac::g_tune[knob] = value;
  return EAC_OK;
}

extern "C" int eac_tune_get(int32_t knob, int32_t* value) {
  if (knob < 0 || knob >= EAC_TUNE_COUNT || !value)
    return eac::fail(EAC_ERR_ARG, "tune: bad knob %d", knob);
  *value = eac::g_tune[knob];
  return EAC_OK;
}
