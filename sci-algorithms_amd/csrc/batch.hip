// Batched GAP-TV / ADMM-TV: many independent measurements of uneven sizes
// solved together.
//
// The reference reconstructs coded frames and measurement blocks one after
// another in Python loops (PnP_SCI/python/pnp_sci_algo.py:498,
// pnp_sci_test_data_realexp_blkproc.py:93).  Here every outer iteration of
// every solve in a batch is one launch of a persistent kernel:
//
//  * each solve keeps its own shapes, buffers and per-iteration launch
//    arguments (FusedArgs, built once per plan and stored on the device);
//  * the (solve, tile) work items of all solves sit in 8 queues, one per XCD:
//    tile bands of each solve are dealt to the XCDs as in the single-solve
//    kernel, so halo re-reads stay in one L2; inside a queue the tiles with
//    the most frames come first (longest-processing-time order);
//  * one workgroup per CU pulls items from its XCD's queue with an atomic
//    counter and steals from the other queues when its own is drained;
//  * interior and border tiles share the launch (the border body is chosen
//    per item, encoded as a negative tile index);
//  * the early-exit decision and repair run per solve after every iteration;
//  * with use_graph the whole run (counter resets, init, every iteration,
//    ADMM's final projection) is captured once as a hipGraph and replayed.

#include "fused.h"

#include <algorithm>
#include <new>
#include <vector>

namespace scigap {

int fill_f64(double* p, int n, double v, hipStream_t s);
int check_params(const scigap_params* p);

namespace {

using namespace fused;

constexpr int kMaxBatchDepth = 6;

// 8-wave tiles of depth 5 are held to 128 VGPRs, two workgroups per CU (they
// would take 132; depth 4 needs 125 on its own).
template <int D, int BY, int MR, int MODE>
__global__ void __launch_bounds__(64 * BY)
__attribute__((amdgpu_waves_per_eu(BY == 8 && D == 5 ? 4 : 1)))
k_fused_batch(const FusedArgs* __restrict__ descs, const int2* __restrict__ queue,
              const int* __restrict__ qoff, int* counters) {
  __shared__ FusedLds<D, BY, MR> sh;
  __shared__ int s_item;
  const int home = blockIdx.x & 7;
  for (int v = 0; v < 8; ++v) {
    const int qi = (home + v) & 7;
    const int beg = qoff[qi], cnt = qoff[qi + 1] - beg;
    if (cnt == 0) continue;
    for (;;) {
      if (threadIdx.x == 0) s_item = atomicAdd(counters + qi, 1);
      __syncthreads();
      const int q = s_item;
      __syncthreads();
      if (q >= cnt) break;
      const int2 it = queue[beg + q];
      if (it.y >= 0)
        fused_tile<D, BY, MR, MODE, true, false>(descs[it.x], it.y, sh);
      else
        fused_tile<D, BY, MR, MODE, false, false>(descs[it.x], ~it.y, sh);
      __syncthreads();
    }
  }
}

// Early-exit decision for every frame of every solve of the batch (one
// workgroup per (solve, frame): grid n x max T); *any counts the stopped
// frames of the whole batch.
template <int D>
__global__ void __launch_bounds__(256) k_decide_batch(const FusedArgs* __restrict__ descs,
                                                      int* any) {
  const FusedArgs& a = descs[blockIdx.x];
  const int t = blockIdx.y;
  if (t >= a.T) return;
  decide_frame<D>(a, t);
  if (threadIdx.x == 0 && a.stop[t] < D) atomicAdd(any, 1);
}

// Early-exit repair over all (solve, tile) items; returns at once when no
// frame of the batch stopped early.
template <int D, int BY, int MR, int MODE>
__global__ void __launch_bounds__(64 * BY) k_fix_batch(const FusedArgs* __restrict__ descs,
                                                      const int2* __restrict__ items, int n,
                                                      const int* any) {
  __shared__ FixLds<D, BY, MR> sh;
  if (*any == 0) return;
  for (int e = blockIdx.x; e < n; e += gridDim.x) {
    const int2 it = items[e];
    const FusedArgs& a = descs[it.x];
    if (*a.nexit == 0) continue;
    if (it.y >= 0)
      fused_fix_tile<D, BY, MR, MODE, true>(a, it.y, sh, a.stop);
    else
      fused_fix_tile<D, BY, MR, MODE, false>(a, ~it.y, sh, a.stop);
    __syncthreads();
  }
}

// Initial residual of every solve (k_init per solve; blockIdx.y = solve).
__global__ void __launch_bounds__(256) k_init_batch(const FusedArgs* __restrict__ d0, float gam) {
  const FusedArgs& a = d0[blockIdx.y];
  const int64_t N = int64_t(a.H) * a.W;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < N;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float yv = a.y[i];
    double sum = 0.0;
    for (int t = 0; t < a.T; ++t) {
      const float f = a.phi[t * N + i];
      const float xv = a.from_y ? yv * f : a.x_in[t * N + i];
      sum = acc_prod(xv, f, sum);
    }
    const float yb = float(sum);
    const float s = a.phi_sum[i];
    float* r = const_cast<float*>(a.r_in);
    if (a.mode == GAP_ACC) {
      const float v1 = yv - yb;
      const_cast<float*>(a.y1_in)[i] = v1;
      r[i] = (v1 - yb) / s;
    } else if (a.mode == GAP_PLAIN) {
      r[i] = (yv - yb) / s;
    } else {
      r[i] = (yv - yb) / (s + gam);
    }
  }
}

// ADMM's last iteration for every solve: x = theta + b + lambda*r*Phi.
__global__ void __launch_bounds__(256) k_admm_final_batch(const FusedArgs* __restrict__ fin) {
  const FusedArgs& a = fin[blockIdx.y];
  const int64_t N = int64_t(a.H) * a.W;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < N;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float lr = a.lam * a.r_in[i];
    const float yv = a.from_y ? a.y[i] : 0.f;
    for (int t = 0; t < a.T; ++t) {
      const size_t o = size_t(t) * N + i;
      const float f = a.phi[o];
      const float th = a.from_y ? yv * f : a.x_in[o];
      a.x_out[o] = fmaf(lr, f, th + (a.b ? a.b[o] : 0.f));
    }
  }
}

__global__ void k_fill_i32_b(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Tile geometry of the batch kernel for (D, BY).
struct Geom {
  int ox, oy, lh, hc;
};
Geom geom_of(int D, int BY) {
  Geom g;
  g.hc = (D + 1) & ~1;
  g.ox = 128 - 2 * g.hc;
  g.lh = BY * D;
  g.oy = g.lh - 2 * D;
  return g;
}

int default_rows(int D) { return D < 4 ? 16 : 8; }

int pick_rows(int D, const scigap_batch_options* opt) {
  if (D == 4 && opt && (opt->tile_rows == 8 || opt->tile_rows == 12 || opt->tile_rows == 16))
    return opt->tile_rows;
  return default_rows(D);
}

// Layout of the batch workspace.
struct BatchLayout {
  std::vector<SolveState> st;
  FusedArgs* descs = nullptr;   // [fused_iters][n]
  FusedArgs* fin = nullptr;     // [n] ADMM final projection
  int2* queue = nullptr;        // per-XCD queues, concatenated
  int* qoff = nullptr;          // [9]
  int2* items = nullptr;        // all items (repair pass)
  int* counters = nullptr;      // [fused_iters][8]
  int* any = nullptr;           // [fused_iters]
  int* zeroed = nullptr;        // zeroed before every run: nexit [K] of every solve
  int zeroed_count = 0;
  int nitems = 0;
  int maxT = 0;
  size_t bytes = 0;
};

int batch_tiles(int64_t H, int64_t W, const Geom& g, int* tiles_x) {
  const int tx = tiles_across(int(W), 128, g.hc, g.ox), ty = int((H + g.oy - 1) / g.oy);
  if (tiles_x) *tiles_x = tx;
  return tx * ty;
}

BatchLayout layout(const scigap_params* p, const scigap_batch_item* items, int n, int BY,
                   char* base) {
  BatchLayout L;
  const int D = p->tv_iter_max - 1;
  const int K = p->iters, KF = K - (p->method == SCIGAP_METHOD_ADMM ? 1 : 0);
  const Geom g = geom_of(D, BY);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* q = base ? base + off : nullptr;
    off += align256(bytes);
    return q;
  };
  int total_tiles = 0;
  size_t zc = 0;
  for (int i = 0; i < n; ++i) zc += size_t(K);
  L.zeroed = reinterpret_cast<int*>(take(zc * 4));
  L.zeroed_count = int(zc);
  size_t zoff = 0;
  L.st.resize(n);
  for (int i = 0; i < n; ++i) {
    const size_t N = size_t(items[i].H) * items[i].W, T = size_t(items[i].T);
    SolveState& s = L.st[i];
    s.xbuf = reinterpret_cast<float*>(take(T * N * 4));
    if (p->method == SCIGAP_METHOD_ADMM) s.bbuf = reinterpret_cast<float*>(take(T * N * 4));
    for (int j = 0; j < 2; ++j) s.r[j] = reinterpret_cast<float*>(take(N * 4));
    if (p->method == SCIGAP_METHOD_GAP && p->accelerate)
      for (int j = 0; j < 2; ++j) s.y1[j] = reinterpret_cast<float*>(take(N * 4));
    s.yb = reinterpret_cast<float*>(take(N * 4));
    s.stop = reinterpret_cast<int*>(take(T * 4));
    s.nexit = L.zeroed ? L.zeroed + zoff : nullptr;
    zoff += size_t(K);
    const int tiles = batch_tiles(items[i].H, items[i].W, g, nullptr);
    s.eslot_stride = tiles;
    s.eslot = reinterpret_cast<float*>(take(T * D * size_t(tiles) * 4));
    total_tiles += tiles;
    L.maxT = std::max(L.maxT, int(T));
  }
  L.descs = reinterpret_cast<FusedArgs*>(take(size_t(KF > 0 ? KF : 1) * n * sizeof(FusedArgs)));
  L.fin = reinterpret_cast<FusedArgs*>(take(size_t(n) * sizeof(FusedArgs)));
  L.queue = reinterpret_cast<int2*>(take(size_t(total_tiles) * sizeof(int2)));
  L.qoff = reinterpret_cast<int*>(take(9 * sizeof(int)));
  L.items = reinterpret_cast<int2*>(take(size_t(total_tiles) * sizeof(int2)));
  L.counters = reinterpret_cast<int*>(take(size_t(KF > 0 ? KF : 1) * 8 * sizeof(int)));
  L.any = reinterpret_cast<int*>(take(size_t(KF > 0 ? KF : 1) * sizeof(int)));
  L.nitems = total_tiles;
  L.bytes = off;
  return L;
}

int check_batch(const scigap_params* p, const scigap_batch_item* items, int n) {
  if (!p || !items) return fail(SCIGAP_ENULL, "batch: null params/items");
  if (n <= 0) return fail(SCIGAP_EINVAL, "batch: n must be > 0");
  scigap_params q = *p;
  for (int i = 0; i < n; ++i) {
    q.H = items[i].H;
    q.W = items[i].W;
    q.T = items[i].T;
    if (int rc = check_params(&q)) return rc;
    if (!items[i].y || !items[i].phi || !items[i].phi_sum || !items[i].x_out)
      return fail(SCIGAP_ENULL, "batch item %d: null input/output", i);
    if ((reinterpret_cast<uintptr_t>(items[i].x_out) | reinterpret_cast<uintptr_t>(items[i].phi) |
         reinterpret_cast<uintptr_t>(items[i].y) | reinterpret_cast<uintptr_t>(items[i].x0)) % 8)
      return fail(SCIGAP_EINVAL, "batch item %d: buffers must be 8-byte aligned", i);
  }
  const int D = p->tv_iter_max - 1;
  if (D < 1 || D > kMaxBatchDepth)
    return fail(SCIGAP_EUNSUPPORTED, "batch: tv_iter_max must be in [2, %d]", kMaxBatchDepth + 1);
  if (p->iters < 1) return fail(SCIGAP_EINVAL, "batch: iters must be >= 1");
  return SCIGAP_OK;
}

}  // namespace

}  // namespace scigap

using namespace scigap;
using namespace scigap::fused;

struct scigap_batch_plan {
  scigap_params p;
  scigap_batch_options opt;
  std::vector<scigap_batch_item> items;
  BatchLayout L;
  int n = 0, D = 0, BY = 16, KF = 0, mode = 0;
  unsigned grid = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

// Workgroups of the persistent kernel that the device holds at once.
template <int D, int BY, int MODE>
unsigned resident_batch_grid() {
  int per_cu = 0, dev = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_fused_batch<D, BY, 1, MODE>, 64 * BY,
                                                   0) != hipSuccess ||
      hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 256;
  return unsigned(std::max(per_cu, 1) * std::max(cus, 1));
}

// Launches every fused iteration of the plan; with P->grid == 0 it only
// sizes the persistent grid (no launch), which plan creation does once.
template <int D, int BY, int MODE>
int enqueue_iters(scigap_batch_plan* P, hipStream_t s) {
  if (P->grid == 0) {
    P->grid = resident_batch_grid<D, BY, MODE>();
    return SCIGAP_OK;
  }
  const BatchLayout& L = P->L;
  const int n = P->n;
  for (int k = 0; k < P->KF; ++k) {
    k_fused_batch<D, BY, 1, MODE><<<P->grid, 64 * BY, 0, s>>>(L.descs + size_t(k) * n, L.queue,
                                                             L.qoff, L.counters + k * 8);
    SCIGAP_LAUNCH_CHECK("batched fused launch");
    k_decide_batch<D><<<dim3(unsigned(n), unsigned(L.maxT)), 256, 0, s>>>(L.descs + size_t(k) * n,
                                                                         L.any + k);
    SCIGAP_LAUNCH_CHECK("batched decision launch");
    k_fix_batch<D, BY, 1, MODE><<<P->grid, 64 * BY, 0, s>>>(L.descs + size_t(k) * n, L.items,
                                                            L.nitems, L.any + k);
    SCIGAP_LAUNCH_CHECK("batched repair launch");
  }
  return SCIGAP_OK;
}

template <int D, int BY>
int enqueue_mode(scigap_batch_plan* P, hipStream_t s) {
  switch (P->mode) {
    case GAP_ACC: return enqueue_iters<D, BY, GAP_ACC>(P, s);
    case GAP_PLAIN: return enqueue_iters<D, BY, GAP_PLAIN>(P, s);
    default: return enqueue_iters<D, BY, ADMM>(P, s);
  }
}

int enqueue_fused(scigap_batch_plan* P, hipStream_t s) {
  switch (P->D) {
    case 1: return enqueue_mode<1, 16>(P, s);
    case 2: return enqueue_mode<2, 16>(P, s);
    case 3: return enqueue_mode<3, 16>(P, s);
    case 4:
      if (P->BY == 8) return enqueue_mode<4, 8>(P, s);
      if (P->BY == 12) return enqueue_mode<4, 12>(P, s);
      return enqueue_mode<4, 16>(P, s);
    case 5: return enqueue_mode<5, 8>(P, s);
    case 6: return enqueue_mode<6, 8>(P, s);
  }
  return fail(SCIGAP_EUNSUPPORTED, "batch depth %d", P->D);
}

int enqueue_all(scigap_batch_plan* P, hipStream_t s) {
  const BatchLayout& L = P->L;
  const int KF = P->KF > 0 ? P->KF : 1;
  k_fill_i32_b<<<(KF * 8 + 255) / 256, 256, 0, s>>>(L.counters, KF * 8, 0);
  SCIGAP_LAUNCH_CHECK("counter reset");
  k_fill_i32_b<<<(KF + 255) / 256, 256, 0, s>>>(L.any, KF, 0);
  SCIGAP_LAUNCH_CHECK("counter reset");
  k_fill_i32_b<<<(L.zeroed_count + 255) / 256, 256, 0, s>>>(L.zeroed, L.zeroed_count, 0);
  SCIGAP_LAUNCH_CHECK("counter reset");
  int64_t maxN = 0;
  for (auto& it : P->items) maxN = std::max<int64_t>(maxN, it.H * it.W);
  const unsigned gx = unsigned(std::min<int64_t>((maxN + 255) / 256, 1024));
  k_init_batch<<<dim3(gx, P->n), 256, 0, s>>>(L.descs, P->p.gamma);
  SCIGAP_LAUNCH_CHECK("batched init launch");
  if (int rc = enqueue_fused(P, s)) return rc;
  if (P->mode == ADMM) {
    k_admm_final_batch<<<dim3(gx, P->n), 256, 0, s>>>(L.fin);
    SCIGAP_LAUNCH_CHECK("batched final launch");
  }
  return SCIGAP_OK;
}

}  // namespace

extern "C" {

size_t scigap_batch_workspace_bytes(const scigap_params* p, const scigap_batch_item* items, int n,
                                    const scigap_batch_options* opt) {
  if (check_batch(p, items, n) != SCIGAP_OK) return 0;
  return layout(p, items, n, pick_rows(p->tv_iter_max - 1, opt), nullptr).bytes;
}

int scigap_batch_plan_create(scigap_batch_plan** out, const scigap_params* p,
                             const scigap_batch_item* items, int n, void* workspace,
                             size_t workspace_bytes, const scigap_batch_options* opt,
                             void* stream) {
  clear_error();
  if (!out) return fail(SCIGAP_ENULL, "plan pointer is NULL");
  *out = nullptr;
  if (int rc = check_batch(p, items, n)) return rc;
  const int D = p->tv_iter_max - 1;
  const int BY = pick_rows(D, opt);
  BatchLayout L = layout(p, items, n, BY, static_cast<char*>(workspace));
  if (!workspace || workspace_bytes < L.bytes)
    return fail(SCIGAP_EWORKSPACE, "batch workspace needs %zu bytes, got %zu", L.bytes,
                workspace_bytes);
  auto* P = new (std::nothrow) scigap_batch_plan;
  if (!P) return fail(SCIGAP_EINVAL, "out of host memory");
  P->p = *p;
  P->opt = opt ? *opt : scigap_batch_options{1, 0, 0, 0};
  P->items.assign(items, items + n);
  P->L = L;
  P->n = n;
  P->D = D;
  P->BY = BY;
  P->mode = solver_mode(*p);
  P->KF = fused_iterations(*p);
  P->grid = unsigned(P->opt.grid > 0 ? P->opt.grid : 0);
  if (P->grid == 0) {
    if (int rc = enqueue_fused(P, nullptr)) {  // sizes the grid, launches nothing
      delete P;
      return rc;
    }
  }

  // Per-iteration launch arguments of every solve.
  const Geom g = geom_of(D, BY);
  std::vector<FusedArgs> descs(size_t(std::max(P->KF, 1)) * n), fin(n);
  for (int i = 0; i < n; ++i) {
    const scigap_batch_item& it = items[i];
    int tiles_x = 0;
    const int ntiles = batch_tiles(it.H, it.W, g, &tiles_x);
    // (ADMM with one iteration runs no fused iteration; its k = 0 arguments
    // still drive the initial residual.)
    for (int k = 0; k < std::max(P->KF, 1); ++k) {
      FusedArgs a = iteration_args(*p, k, it.H, it.W, it.T, it.y, it.phi, it.phi_sum, it.x0,
                                   it.x_out, nullptr, nullptr, nullptr, L.st[i]);
      a.tiles_x = tiles_x;
      a.ntiles = ntiles;
      a.per_xcd = (ntiles + 7) / 8;
      a.eslot_tiles = ntiles;
      descs[size_t(k) * n + i] = a;
    }
    if (P->mode == ADMM) {
      const int K = p->iters;
      FusedArgs f{};
      // One iteration: x = X_0 + lambda*r_0*Phi with X_0 = x0 or At(y), b = 0.
      f.from_y = (K == 1 && !it.x0) ? 1 : 0;
      f.x_in = K == 1 ? it.x0 : state_buffer(*p, L.st[i], it.x_out, K - 1);
      f.x_out = it.x_out;
      f.b = K == 1 ? nullptr : L.st[i].bbuf;
      f.r_in = L.st[i].r[(K - 1) & 1];
      f.phi = it.phi;
      f.y = it.y;
      f.H = int(it.H);
      f.W = int(it.W);
      f.T = int(it.T);
      f.lam = p->lambda_;
      fin[i] = f;
    }
  }

  // Per-XCD queues: band b of every solve goes to queue b; solves with the
  // most frames first (or in submission order).
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  if (P->opt.order == 0)
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      if (items[a].T != items[b].T) return items[a].T > items[b].T;
      return items[a].H * items[a].W > items[b].H * items[b].W;
    });
  std::vector<int2> queue, all;
  std::vector<int> qoff(9, 0);
  for (int b = 0; b < 8; ++b) {
    qoff[b] = int(queue.size());
    for (int i : order) {
      const scigap_batch_item& it = items[i];
      int tiles_x = 0;
      const int ntiles = batch_tiles(it.H, it.W, g, &tiles_x);
      const int per = (ntiles + 7) / 8;
      for (int t = b * per; t < std::min(ntiles, (b + 1) * per); ++t) {
        bool inside;
        switch (D) {
          case 1: inside = tile_inside<1, 16, 1>(t, tiles_x, int(it.H), int(it.W)); break;
          case 2: inside = tile_inside<2, 16, 1>(t, tiles_x, int(it.H), int(it.W)); break;
          case 3: inside = tile_inside<3, 16, 1>(t, tiles_x, int(it.H), int(it.W)); break;
          case 4:
            inside = BY == 8    ? tile_inside<4, 8, 1>(t, tiles_x, int(it.H), int(it.W))
                     : BY == 12 ? tile_inside<4, 12, 1>(t, tiles_x, int(it.H), int(it.W))
                                : tile_inside<4, 16, 1>(t, tiles_x, int(it.H), int(it.W));
            break;
          case 5: inside = tile_inside<5, 8, 1>(t, tiles_x, int(it.H), int(it.W)); break;
          default: inside = tile_inside<6, 8, 1>(t, tiles_x, int(it.H), int(it.W)); break;
        }
        queue.push_back(make_int2(i, inside ? t : ~t));
      }
    }
  }
  qoff[8] = int(queue.size());
  all = queue;

  hipStream_t s = as_stream(stream);
  auto up = [&](void* dst, const void* src, size_t bytes) {
    return bytes ? hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) : hipSuccess;
  };
  hipError_t e = hipSuccess;
  e = up(L.descs, descs.data(), descs.size() * sizeof(FusedArgs));
  if (e == hipSuccess && P->mode == ADMM) e = up(L.fin, fin.data(), fin.size() * sizeof(FusedArgs));
  if (e == hipSuccess) e = up(L.queue, queue.data(), queue.size() * sizeof(int2));
  if (e == hipSuccess) e = up(L.qoff, qoff.data(), qoff.size() * sizeof(int));
  if (e == hipSuccess) e = up(L.items, all.data(), all.size() * sizeof(int2));
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // host vectors go out of scope
  if (e != hipSuccess) {
    delete P;
    return hip_fail(e, "batch plan upload");
  }
  *out = P;
  return SCIGAP_OK;
}

int scigap_batch_run(scigap_batch_plan* P, void* stream) {
  clear_error();
  if (!P) return fail(SCIGAP_ENULL, "plan is NULL");
  hipStream_t s = as_stream(stream);
  if (!P->opt.use_graph) return enqueue_all(P, s);
  if (!P->exec) {
    // Capture on a private stream: the caller's stream may be the legacy
    // null stream, which cannot be captured.  Capturing executes nothing.
    hipStream_t cs = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "capture stream");
    e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      hipStreamDestroy(cs);
      return hip_fail(e, "begin graph capture");
    }
    int rc = enqueue_all(P, cs);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    hipStreamDestroy(cs);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "end graph capture");
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      hipGraphDestroy(g);
      return hip_fail(e, "graph instantiate");
    }
    P->graph = g;
    P->exec = ex;
  }
  hipError_t e = hipGraphLaunch(P->exec, s);
  return e == hipSuccess ? SCIGAP_OK : hip_fail(e, "graph launch");
}

int scigap_batch_plan_destroy(scigap_batch_plan* P) {
  clear_error();
  if (!P) return SCIGAP_OK;
  if (P->exec) hipGraphExecDestroy(P->exec);
  if (P->graph) hipGraphDestroy(P->graph);
  delete P;
  return SCIGAP_OK;
}

}  // extern "C"
