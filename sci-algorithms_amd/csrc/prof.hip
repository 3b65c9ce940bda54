// Per-kernel timing with HIP events on the launching stream.
//
// While enabled, ProfScope records an event before and after a launch; the
// pairs are resolved when profiling ends.  Events are pooled and reused.

#include "scigap_internal.h"

#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace scigap {

namespace {

struct Rec {
  std::string name;
  hipEvent_t a, b;
};

std::mutex g_mu;
std::atomic<bool> g_on{false};  // read without the lock on every launch
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

}  // namespace

bool profiling() { return g_on; }

ProfScope::ProfScope(const char* name, hipStream_t s) : name_(name), s_(s) {
  if (!g_on) return;
  std::lock_guard<std::mutex> lk(g_mu);
  a_ = take_event();
  b_ = take_event();
  if (a_ && b_) hipEventRecord(static_cast<hipEvent_t>(a_), s_);
}

ProfScope::~ProfScope() {
  if (!a_ || !b_) return;
  std::lock_guard<std::mutex> lk(g_mu);
  hipEventRecord(static_cast<hipEvent_t>(b_), s_);
  g_recs.push_back({name_, static_cast<hipEvent_t>(a_), static_cast<hipEvent_t>(b_)});
}

}  // namespace scigap

using namespace scigap;

extern "C" {

int scigap_profile_begin(void) {
  clear_error();
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto& r : g_recs) {
    g_pool.push_back(r.a);
    g_pool.push_back(r.b);
  }
  g_recs.clear();
  g_on = true;
  return SCIGAP_OK;
}

int scigap_profile_end(scigap_kernel_stat* out, int max, int* n) {
  clear_error();
  std::lock_guard<std::mutex> lk(g_mu);
  g_on = false;
  std::vector<scigap_kernel_stat> stats;
  for (auto& r : g_recs) {
    hipError_t e = hipEventSynchronize(r.b);
    if (e != hipSuccess) return hip_fail(e, "profile event sync");
    float ms = 0.f;
    e = hipEventElapsedTime(&ms, r.a, r.b);
    if (e != hipSuccess) return hip_fail(e, "profile elapsed time");
    scigap_kernel_stat* st = nullptr;
    for (auto& q : stats)
      if (r.name == q.name) st = &q;
    if (!st) {
      stats.push_back({});
      st = &stats.back();
      std::strncpy(st->name, r.name.c_str(), sizeof(st->name) - 1);
    }
    st->launches += 1;
    st->total_ms += ms;
    g_pool.push_back(r.a);
    g_pool.push_back(r.b);
  }
  g_recs.clear();
  if (n) *n = int(stats.size());
  for (int i = 0; i < int(stats.size()) && i < max; ++i) out[i] = stats[i];
  return SCIGAP_OK;
}

}  // extern "C"
