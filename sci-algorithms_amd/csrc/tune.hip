// Autotuner: picks the fused kernel's tile height and engine for a solve
// shape, and the TV fusion depth (dual updates per launch) for deep TV, by
// timing the candidates with HIP events on the caller's stream and data.
//
// Results live in a process-wide cache keyed by the problem shape (the only
// global state of the library, guarded by a mutex).  When the environment
// variable SCIGAP_TUNE_CACHE names a file, entries are loaded from it on
// first use and every new entry is appended to it, so later processes start
// tuned.  scigap_solve uses a cached entry when its tile_rows and engine are
// both 0; scigap_tv_chambolle uses the cached depth for depth 0.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "scigap_internal.h"

namespace scigap {

int max_chunk_depth();
int check_params(const scigap_params* p);

namespace {

// kind 0: solve (a = method*2 + accelerate, b = tv_iter_max); kind 1: TV (b = n).
using Key = std::tuple<int, int64_t, int64_t, int64_t, int, int>;
struct Entry {
  int tile_rows = 0, engine = 0, depth = 0;
  double ms = 0.0;
};

std::mutex g_mu;
std::map<Key, Entry> g_cache;
bool g_loaded = false;

const char* cache_path() {
  const char* p = std::getenv("SCIGAP_TUNE_CACHE");
  return (p && *p) ? p : nullptr;
}

// Caller holds g_mu.
void load_locked() {
  if (g_loaded) return;
  g_loaded = true;
  const char* path = cache_path();
  if (!path) return;
  FILE* f = std::fopen(path, "r");
  if (!f) return;
  int kind, a, b, rows, eng, depth;
  long long H, W, T;
  double ms;
  while (std::fscanf(f, "%d %lld %lld %lld %d %d %d %d %d %lf", &kind, &H, &W, &T, &a, &b, &rows,
                     &eng, &depth, &ms) == 10) {
    Entry e;
    e.tile_rows = rows;
    e.engine = eng;
    e.depth = depth;
    e.ms = ms;
    g_cache[Key(kind, H, W, T, a, b)] = e;
  }
  std::fclose(f);
}

void store(const Key& k, const Entry& e) {
  std::lock_guard<std::mutex> lock(g_mu);
  load_locked();
  g_cache[k] = e;
  if (const char* path = cache_path()) {
    if (FILE* f = std::fopen(path, "a")) {
      std::fprintf(f, "%d %lld %lld %lld %d %d %d %d %d %.6f\n", std::get<0>(k),
                   (long long)std::get<1>(k), (long long)std::get<2>(k),
                   (long long)std::get<3>(k), std::get<4>(k), std::get<5>(k), e.tile_rows,
                   e.engine, e.depth, e.ms);
      std::fclose(f);
    }
  }
}

bool lookup(const Key& k, Entry* e) {
  std::lock_guard<std::mutex> lock(g_mu);
  load_locked();
  auto it = g_cache.find(k);
  if (it == g_cache.end()) return false;
  *e = it->second;
  return true;
}

Key solve_key(const scigap_params& p) {
  return Key(0, p.H, p.W, p.T, p.method * 2 + (p.accelerate ? 1 : 0), p.tv_iter_max);
}

// Milliseconds per call of fn (the best of `reps` after one warm-up), or a
// negative status.
template <class F>
double time_on_stream(F&& fn, hipStream_t s, int reps, int* rc) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) {
    *rc = fail(SCIGAP_EHIP, "hipEventCreate failed");
    return -1.0;
  }
  if (hipEventCreate(&e1) != hipSuccess) {
    hipEventDestroy(e0);
    *rc = fail(SCIGAP_EHIP, "hipEventCreate failed");
    return -1.0;
  }
  double best = -1.0;
  *rc = fn();
  for (int r = 0; r < reps && *rc == SCIGAP_OK; ++r) {
    hipEventRecord(e0, s);
    *rc = fn();
    hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) {
      *rc = fail(SCIGAP_EHIP, "hipEventSynchronize failed");
      break;
    }
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (best < 0.0 || ms < best) best = ms;
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return best;
}

}  // namespace

// Tuned TV fusion depth for a deep TV (n - 1 > 6 updates); 6 when untuned.
int tuned_tv_depth(int64_t H, int64_t W, int64_t T, int n) {
  Entry e;
  if (lookup(Key(1, H, W, T, 0, n), &e) && e.depth >= 1 && e.depth <= max_chunk_depth())
    return e.depth;
  return max_chunk_depth();
}

// Tuned (tile_rows, engine) of a solve shape, if any.
bool tuned_solve_config(const scigap_params& p, int* tile_rows, int* engine) {
  Entry e;
  if (!lookup(solve_key(p), &e)) return false;
  *tile_rows = e.tile_rows;
  *engine = e.engine;
  return true;
}

}  // namespace scigap

using namespace scigap;

extern "C" {

int scigap_autotune_solve(const scigap_params* p, const float* y, const float* phi,
                          const float* phi_sum, float* x_scratch, int iters, void* workspace,
                          size_t workspace_bytes, void* stream, scigap_tune_result* res) {
  clear_error();
  if (int rc = check_params(p)) return rc;
  if (!res) return fail(SCIGAP_ENULL, "result is NULL");
  if (iters < 1) return fail(SCIGAP_EINVAL, "iters must be >= 1");
  hipStream_t s = as_stream(stream);
  scigap_params q = *p;
  q.iters = iters;
  std::vector<std::pair<int, int>> cands;
  if (p->tv_iter_max == 5) {
    for (int rows : {8, 12, 16})
      for (int eng : {1, 2, 3, 4, 5, 6}) cands.emplace_back(rows, eng);
  } else {
    cands.emplace_back(0, 1);
    cands.emplace_back(0, 3);
    cands.emplace_back(0, 6);
  }
  Entry best;
  double def_ms = -1.0;
  int n = 0;
  for (auto [rows, eng] : cands) {
    q.tile_rows = rows;
    q.engine = eng;
    int rc = SCIGAP_OK;
    const double ms = time_on_stream(
        [&]() {
          return scigap_solve(&q, y, phi, phi_sum, nullptr, x_scratch, nullptr, nullptr, nullptr,
                              workspace, workspace_bytes, stream);
        },
        s, 2, &rc);
    if (rc != SCIGAP_OK) return rc;
    ++n;
    if (n == 1) def_ms = ms;
    if (best.ms == 0.0 || ms < best.ms) {
      best.tile_rows = rows;
      best.engine = eng;
      best.ms = ms;
    }
  }
  store(solve_key(*p), best);
  res->tile_rows = best.tile_rows;
  res->engine = best.engine;
  res->tv_depth = 0;
  res->candidates = n;
  res->best_ms = best.ms / iters;
  res->default_ms = def_ms / iters;
  return SCIGAP_OK;
}

int scigap_autotune_tv(const float* image, float* out, int64_t H, int64_t W, int64_t T,
                       float weight, float eps, int n_iter_max, void* workspace,
                       size_t workspace_bytes, void* stream, scigap_tune_result* res) {
  clear_error();
  if (!res) return fail(SCIGAP_ENULL, "result is NULL");
  if (n_iter_max < 2) return fail(SCIGAP_EINVAL, "n_iter_max must be >= 2 to tune");
  hipStream_t s = as_stream(stream);
  const int D = n_iter_max - 1;
  Entry best;
  double def_ms = -1.0;
  int n = 0;
  const int dmax = max_chunk_depth();
  for (int d = dmax; d >= 1; --d) {
    if (d > D) continue;
    int rc = SCIGAP_OK;
    const double ms = time_on_stream(
        [&]() {
          return scigap_tv_chambolle_depth(image, out, H, W, T, weight, eps, n_iter_max, d,
                                           nullptr, workspace, workspace_bytes, stream);
        },
        s, 2, &rc);
    if (rc != SCIGAP_OK) return rc;
    ++n;
    if (n == 1) def_ms = ms;
    if (best.ms == 0.0 || ms < best.ms) {
      best.depth = d;
      best.ms = ms;
    }
  }
  store(Key(1, H, W, T, 0, n_iter_max), best);
  res->tile_rows = 0;
  res->engine = 0;
  res->tv_depth = best.depth;
  res->candidates = n;
  res->best_ms = best.ms;
  res->default_ms = def_ms;
  return SCIGAP_OK;
}

int scigap_tune_clear(void) {
  std::lock_guard<std::mutex> lock(g_mu);
  g_cache.clear();
  g_loaded = true;  // do not reload the file after an explicit clear
  return SCIGAP_OK;
}

}  // extern "C"
