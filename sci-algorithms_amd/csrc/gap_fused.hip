// Fused GAP-TV / ADMM-TV outer iteration for gfx950.
//
// One launch performs one outer iteration of the reference solvers
// (PnP_SCI/python/pnp_sci_algo.py:638-650 for GAP, :805-836 for ADMM) with the
// scikit-image Chambolle TV (n_iter_max = D+1, i.e. D dual updates) fused in:
//
//   prologue  img_t = x_t + lambda*r*Phi_t     (GAP: x + lambda*At(r))
//                      theta_t + lambda*r*Phi_t (ADMM: v = x - b, see below)
//   TV        D Chambolle updates of img_t on a register tile (tv_tile.h)
//   epilogue  yb += Phi_t * x_t^{new} over the frames, then the residual of
//             the NEXT iteration: r' = (y1' - yb)/Phi_sum with
//             y1' = y1 + (y - yb) (accelerated GAP), r' = (y - yb)/Phi_sum
//             (plain GAP) or r' = (y - yb)/(Phi_sum + gamma) (ADMM).
//
// Carrying r (one plane) from one launch to the next replaces the separate
// A(x) pass of the reference, so an iteration reads x, Phi (T planes each),
// r, y, y1, Phi_sum and writes x (T planes), r, y1: (3T+6)*4 bytes per pixel.
//
// ADMM algebra.  With v = x - b the reference's updates reduce to
//   v_k = theta_k + lambda*r_k*Phi,  theta_{k+1} = TV(v_k),
//   b_{k+1} = theta_{k+1} - v_k,    theta_{k+1} + b_{k+1} = 2 theta_{k+1} - v_k,
// so b never has to be read inside the loop.  The returned x of the last
// iteration is theta + b + lambda*r*Phi (k_admm_final); b is written by the
// iteration before it (and every iteration when per-iteration errors are
// requested, since x_k = v_k + b_k).
//
// Early exit.  Each launch stores per-tile partial energies E_0..E_{D-1} of
// every frame; a decision kernel (one workgroup per frame) sums them in a
// fixed order and derives each frame's exit iteration, and a repair kernel
// (which returns at once when no frame stopped early) recomputes the frames
// that stopped at s < D from the unchanged inputs (x, r and y1 are
// double-buffered) and corrects yb and the epilogue.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <mutex>

#include "fused.h"

namespace scigap {

int fill_f64(double* p, int n, double v, hipStream_t s);
size_t tv_chunked_workspace(int64_t H, int64_t W, int64_t T, int n, int nchunks_max);
int tv_chunked(const float* image, float* out, int64_t H, int64_t W, int64_t T, float weight,
               float eps, int n, int depth, int* stop_iter, void* ws, hipStream_t s);
int tuned_tv_depth(int64_t H, int64_t W, int64_t T, int n);
bool tuned_solve_config(const scigap_params& p, int* tile_rows, int* engine);

namespace {

using namespace fused;

constexpr int kMaxFused = 6;

// Workgroup shape per fused depth: 16 waves (1024 lanes) of one D-row strip
// each while D <= 4, 8 waves beyond (the strips get taller).  Load tile
// 128 x BY*D, interior (128-2*HC) x (BY*D-2D).
template <int D> struct FusedShape {
  static constexpr int BY = D <= 4 ? 16 : 8;
  static constexpr int MR = 1;
};

// Repair-kernel tiles (independent of the main pass' tiles: it recomputes
// whole frames): for D = 4 the 8-wave shape of the default main pass, so
// that two workgroups share a CU.
template <int D> struct FixShape {
  static constexpr int BY = D == 4 ? 8 : FusedShape<D>::BY;
};

// Tile grid of the (D, BY) geometry written into the launch arguments.
template <int D, int BY, int MR>
unsigned set_tiles(FusedArgs& a) {
  using G = TileGeom<D, BY, MR>;
  a.tiles_x = tiles_x_of<D, BY, MR>(a.W);
  const int tiles_y = (a.H + G::OY - 1) / G::OY;
  a.ntiles = a.tiles_x * tiles_y;
  a.per_xcd = (a.ntiles + 7) / 8;
  return unsigned(a.per_xcd * 8);
}

template <int D, int BY, int MODE, bool IQA, bool X1, int PART, bool DMA = false, int MR = 1>
void launch_kernel(const FusedArgs& a, unsigned grid, hipStream_t s) {
  k_fused<D, BY, MR, MODE, IQA, X1, PART, DMA><<<grid, 64 * BY, 0, s>>>(a);
}

// Engines of the fused kernel (scigap_params.engine):
//   1  two LDS exchanges per TV update, interior and border tiles in one launch
//   2  one exchange per update (tv_run_x1), one launch
//   3  as 1, interior and border tiles in separate launches
//   4  as 2, separate launches
//   5  as 3, and interior tiles prefetch the next frame straight into LDS
//      (buffer_load ... lds) instead of registers (needs W % 4 == 0, else 3)
//   6  as 1, with each tile's frames split over F workgroups (their
//      partial yb summed by k_part_epilogue); F is chosen from the tile
//      count and the number of resident workgroups so that the last round
//      of workgroups is full (large images) or the GPU fills at all (small
//      images)
// The alternatives 2-5 exist for tv_iter_max = 5 (D = 4).  engine 0 is 6.
constexpr int kDefaultEngine = 6;
constexpr int kMaxParts = 8;  // frame parts (yb_part planes in the workspace)

// Workgroups of the fused kernel <D, BY, MODE, IQA> resident on the whole
// device (queried once per instantiation; every device of a process is
// assumed to be the same model).
template <int D, int BY, int MODE, bool IQA, int MR = 1>
int resident_workgroups() {
  static std::atomic<int> cached{0};  // host threads may race to fill it (same value)
  int v = cached.load(std::memory_order_relaxed);
  if (v == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, k_fused<D, BY, MR, MODE, IQA, false, 0>, 64 * BY, 0) != hipSuccess ||
        cus <= 0 || per_cu <= 0)
      return 256;
    v = cus * per_cu;
    cached.store(v, std::memory_order_relaxed);
  }
  return v;
}

// Frame parts per tile: minimise rounds(F) * (frames per part + overhead),
// with rounds(F) = ceil(tiles*F / resident workgroups).
int pick_parts(int tiles, int T, int slots) {
  int best = 1;
  double best_cost = 1e300;
  for (int F = 1; F <= std::min(T, kMaxParts); ++F) {
    const double rounds = double((int64_t(tiles) * F + slots - 1) / slots);
    const double cost = rounds * (double((T + F - 1) / F) + 0.25);
    if (cost < best_cost - 1e-9) {
      best_cost = cost;
      best = F;
    }
  }
  return best;
}

// With `fold_parts`, a main pass split into frame parts leaves the parts'
// epilogue to the caller's decision launch (k_decide_epilogue) and reports
// the split there; otherwise it runs k_part_epilogue itself.
template <int D, int BY, int MODE, bool IQA, int MR = 1>
int launch_main(const FusedArgs& a0, hipStream_t s, int engine, int* fold_parts = nullptr) {
  FusedArgs a = a0;
  const unsigned grid = set_tiles<D, BY, MR>(a);
  if (engine == 0) engine = kDefaultEngine;
  int parts = 1;
  if (engine == 6 && (MODE == TV_ONLY || a.yb_part)) {
    const int slots = resident_workgroups<D, BY, MODE, IQA, MR>();
    parts = pick_parts(a.ntiles, a.T, slots);
  }
  if (parts > 1) {
    a.fsplit = parts;
    a.ntiles *= parts;
    a.per_xcd = (a.ntiles + 7) / 8;
    ProfScope ps(MODE == TV_ONLY ? "tv_fused" : "fused_iteration", s);
    launch_kernel<D, BY, MODE, IQA, false, 0, false, MR>(a, unsigned(a.per_xcd * 8), s);
    if constexpr (MODE != TV_ONLY) {
      if (fold_parts) {
        *fold_parts = parts;
      } else {
        const int64_t N = int64_t(a.H) * a.W;
        const unsigned g = unsigned(std::min<int64_t>((N + 255) / 256, 4096));
        k_part_epilogue<MODE><<<g, 256, 0, s>>>(a);
      }
    }
  } else {
    ProfScope ps(MODE == TV_ONLY ? "tv_fused" : "fused_iteration", s);
    if constexpr (D == 4) {
      constexpr bool x1_fits = sizeof(FusedLds<D, BY, 1, true>) <= 160 * 1024;
      const bool split = engine == 3 || engine == 4 || engine == 5;
      if (engine == 5 && a.W % 4 == 0) {
        launch_kernel<D, BY, MODE, IQA, false, 1, true>(a, grid, s);
        launch_kernel<D, BY, MODE, IQA, false, 2>(a, grid, s);
      } else if (x1_fits && (engine == 2 || engine == 4)) {
        if constexpr (x1_fits) {
          if (split) {
            launch_kernel<D, BY, MODE, IQA, true, 1>(a, grid, s);
            launch_kernel<D, BY, MODE, IQA, true, 2>(a, grid, s);
          } else {
            launch_kernel<D, BY, MODE, IQA, true, 0>(a, grid, s);
          }
        }
      } else if (split) {
        launch_kernel<D, BY, MODE, IQA, false, 1>(a, grid, s);
        launch_kernel<D, BY, MODE, IQA, false, 2>(a, grid, s);
      } else {
        launch_kernel<D, BY, MODE, IQA, false, 0>(a, grid, s);
      }
    } else {
      launch_kernel<D, BY, MODE, IQA, false, 0>(a, grid, s);
    }
  }
  SCIGAP_LAUNCH_CHECK("fused iteration launch");
  return SCIGAP_OK;
}

template <int D, int BY, int MR = 1>
int launch_mode(const FusedArgs& a, hipStream_t s, int engine, int* fold_parts = nullptr) {
  const bool iqa = a.truth != nullptr;
  if (a.mode == GAP_ACC)
    return iqa ? launch_main<D, BY, GAP_ACC, true, MR>(a, s, engine, fold_parts)
               : launch_main<D, BY, GAP_ACC, false, MR>(a, s, engine, fold_parts);
  if (a.mode == GAP_PLAIN)
    return iqa ? launch_main<D, BY, GAP_PLAIN, true, MR>(a, s, engine, fold_parts)
               : launch_main<D, BY, GAP_PLAIN, false, MR>(a, s, engine, fold_parts);
  if (a.mode == ADMM)
    return iqa ? launch_main<D, BY, ADMM, true, MR>(a, s, engine, fold_parts)
               : launch_main<D, BY, ADMM, false, MR>(a, s, engine, fold_parts);
  return launch_main<D, BY, TV_ONLY, false, MR>(a, s, engine, fold_parts);
}

// Tiles of the main launch for (D, BY, MR): the energy slots' row length.
template <int D, int BY, int MR = 1>
int main_tiles(int H, int W) {
  using G = TileGeom<D, BY, MR>;
  return tiles_x_of<D, BY, MR>(W) * ((H + G::OY - 1) / G::OY);
}

template <int D>
int main_tiles_for(int H, int W, int tile_rows, int engine) {
  if constexpr (D == 4) {
    return tile_rows == 12 ? main_tiles<4, 12>(H, W)
           : tile_rows == 16 ? main_tiles<4, 16>(H, W) : main_tiles<4, 8>(H, W);
  } else {
    return main_tiles<D, FusedShape<D>::BY>(H, W);
  }
}

template <int D>
int launch_fused(const FusedArgs& a00, hipStream_t s, int tile_rows, int engine) {
  constexpr int BY = FusedShape<D>::BY, MR = FusedShape<D>::MR;
  FusedArgs a0 = a00;
  if (a0.eslot) {
    a0.eslot_tiles = main_tiles_for<D>(a0.H, a0.W, tile_rows, engine);
    if (a0.eslot_tiles > a0.eslot_stride)
      return fail(SCIGAP_EWORKSPACE, "energy slots hold %d tiles, the launch has %d",
                  a0.eslot_stride, a0.eslot_tiles);
  }
  // When a decision follows, a main pass split into frame parts leaves its
  // parts' epilogue to the decision launch (parts > 1 then).
  int parts = 1;
  int* const fold = a0.stop && a0.eslot ? &parts : nullptr;
  int rc;
  if constexpr (D == 4) {
    // 8 waves measured fastest at 2048x2048x20 (419 us against 509 us for
    // 12 and 478 us for 16 per outer iteration): at about 110 VGPRs two
    // 8-wave workgroups share a CU, so one hides the other's barriers.
    rc = tile_rows == 12   ? launch_mode<4, 12>(a0, s, engine, fold)
         : tile_rows == 16 ? launch_mode<4, 16>(a0, s, engine, fold)
                           : launch_mode<4, 8>(a0, s, engine, fold);
  } else {
    rc = launch_mode<D, BY>(a0, s, engine, fold);
  }
  if (rc) return rc;
  // No stop array: the caller runs without the energy test (eps = 0), so no
  // frame can stop early and there is nothing to decide or repair.
  if (!a0.stop) return SCIGAP_OK;
  if (!a0.eslot) return fail(SCIGAP_EINVAL, "early exit without energy slots");
  FusedArgs a = a0;
  constexpr int FBY = FixShape<D>::BY;
  const unsigned grid = set_tiles<D, FBY, MR>(a);
  {
    ProfScope ps("fused_decide", s);
    if (parts > 1) {
      FusedArgs e = a0;
      e.fsplit = parts;
      const int64_t N = int64_t(a.H) * a.W;
      const unsigned g =
          unsigned(std::min<int64_t>((N + kDecideThreads - 1) / kDecideThreads, 1024));
      const unsigned blocks = unsigned(a.T) + g;
      switch (a.mode) {
        case GAP_ACC: k_decide_epilogue<D, GAP_ACC><<<blocks, kDecideThreads, 0, s>>>(e); break;
        case GAP_PLAIN:
          k_decide_epilogue<D, GAP_PLAIN><<<blocks, kDecideThreads, 0, s>>>(e);
          break;
        default: k_decide_epilogue<D, ADMM><<<blocks, kDecideThreads, 0, s>>>(e); break;
      }
    } else {
      k_decide<D><<<unsigned(a.T), kDecideThreads, 0, s>>>(a);
    }
    SCIGAP_LAUNCH_CHECK("early-exit decision launch");
  }
  {
    ProfScope ps("fused_fix", s);
    switch (a.mode) {
      case GAP_ACC: k_fused_fix<D, FBY, MR, GAP_ACC><<<grid, 64 * FBY, 0, s>>>(a); break;
      case GAP_PLAIN: k_fused_fix<D, FBY, MR, GAP_PLAIN><<<grid, 64 * FBY, 0, s>>>(a); break;
      case ADMM: k_fused_fix<D, FBY, MR, ADMM><<<grid, 64 * FBY, 0, s>>>(a); break;
      default: k_fused_fix<D, FBY, MR, TV_ONLY><<<grid, 64 * FBY, 0, s>>>(a); break;
    }
  }
  SCIGAP_LAUNCH_CHECK("fused early-exit launch");
  return SCIGAP_OK;
}

int launch_fused_depth(int D, const FusedArgs& a, hipStream_t s, int tile_rows = 0,
                       int engine = 0) {
  switch (D) {
    case 1: return launch_fused<1>(a, s, tile_rows, engine);
    case 2: return launch_fused<2>(a, s, tile_rows, engine);
    case 3: return launch_fused<3>(a, s, tile_rows, engine);
    case 4: return launch_fused<4>(a, s, tile_rows, engine);
    case 5: return launch_fused<5>(a, s, tile_rows, engine);
    case 6: return launch_fused<6>(a, s, tile_rows, engine);
  }
  return fail(SCIGAP_EUNSUPPORTED, "fused depth %d not instantiated", D);
}

// Longest energy-slot row any main launch of depth D can need at HxW (the
// tile heights selectable for D = 4 included).
int max_slot_tiles(int D, int H, int W) {
  switch (D) {
    case 1: return main_tiles<1, FusedShape<1>::BY>(H, W);
    case 2: return main_tiles<2, FusedShape<2>::BY>(H, W);
    case 3: return main_tiles<3, FusedShape<3>::BY>(H, W);
    case 4:
      return std::max(main_tiles<4, 8>(H, W),
                      std::max(main_tiles<4, 12>(H, W), main_tiles<4, 16>(H, W)));
    case 5: return main_tiles<5, FusedShape<5>::BY>(H, W);
    case 6: return main_tiles<6, FusedShape<6>::BY>(H, W);
  }
  return 0;
}

unsigned grid_pixels(int64_t N) {
  int64_t b = (N + 255) / 256;
  return unsigned(b < 4096 ? (b < 1 ? 1 : b) : 4096);
}

// Workspace of one solve: SolveState plus the deep-TV buffers.
struct Workspace {
  SolveState st;
  float* vbuf = nullptr;  // T*N TV input (deep-TV path)
  void* tvws = nullptr;   // general TV workspace (deep-TV path)
  size_t bytes = 0;
};

Workspace carve(const scigap_params* p, char* base) {
  Workspace w;
  const size_t N = size_t(p->H) * p->W, T = size_t(p->T);
  const int D = p->tv_iter_max - 1 > 0 ? p->tv_iter_max - 1 : 1;
  const size_t K = size_t(p->iters > 0 ? p->iters : 1);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* q = base ? base + off : nullptr;
    off += align256(bytes);
    return q;
  };
  w.st.xbuf = reinterpret_cast<float*>(take(T * N * 4));
  if (p->method == SCIGAP_METHOD_ADMM) w.st.bbuf = reinterpret_cast<float*>(take(T * N * 4));
  for (int i = 0; i < 2; ++i) w.st.r[i] = reinterpret_cast<float*>(take(N * 4));
  if (p->method == SCIGAP_METHOD_GAP && p->accelerate)
    for (int i = 0; i < 2; ++i) w.st.y1[i] = reinterpret_cast<float*>(take(N * 4));
  if (p->tv_iter_max - 1 >= 1 && p->tv_iter_max - 1 <= kMaxFused) {
    w.st.eslot_stride = max_slot_tiles(D, int(p->H), int(p->W));
    w.st.eslot = reinterpret_cast<float*>(take(T * D * size_t(w.st.eslot_stride) * 4));
  }
  w.st.yb = reinterpret_cast<float*>(take(N * 4));
  w.st.yb_part = reinterpret_cast<float*>(take(std::min<size_t>(T, kMaxParts) * N * 4));
  w.st.stop = reinterpret_cast<int*>(take(T * 4));
  w.st.nexit = reinterpret_cast<int*>(take(K * 4));
  if (p->tv_iter_max - 1 > kMaxFused) {
    w.vbuf = reinterpret_cast<float*>(take(T * N * 4));
    w.tvws = take(tv_chunked_workspace(p->H, p->W, p->T, p->tv_iter_max, p->tv_iter_max));
  }
  w.bytes = off;
  return w;
}

}  // namespace

int check_params(const scigap_params* p) {
  if (!p) return fail(SCIGAP_ENULL, "params is NULL");
  if (int rc = check_shape(p->H, p->W, p->T)) return rc;
  if (p->method != SCIGAP_METHOD_GAP && p->method != SCIGAP_METHOD_ADMM)
    return fail(SCIGAP_EINVAL, "unknown method %d", p->method);
  if (p->iters < 0 || p->iters > (1 << 24))
    return fail(SCIGAP_EINVAL, "iters must be in [0, 2^24]");
  if (p->tv_iter_max < 1) return fail(SCIGAP_EINVAL, "tv_iter_max must be >= 1");
  if (!(p->tv_weight > 0.f) || !std::isfinite(p->tv_weight))
    return fail(SCIGAP_EINVAL, "tv_weight must be finite and > 0");
  // Scalars are checked here, on the host; non-finite values inside the
  // device arrays propagate through the arithmetic as in the reference.
  if (!std::isfinite(p->lambda_) || !std::isfinite(p->gamma) || !std::isfinite(p->tv_eps) ||
      !std::isfinite(p->tv_weight_decay) || !std::isfinite(p->gamma_decay))
    return fail(SCIGAP_EINVAL, "lambda_, gamma, tv_eps and the decay factors must be finite");
  if (p->tile_rows != 0 && p->tile_rows != 8 && p->tile_rows != 12 && p->tile_rows != 16)
    return fail(SCIGAP_EINVAL, "tile_rows must be 0, 8, 12 or 16");
  if (p->engine < 0 || p->engine > 6) return fail(SCIGAP_EINVAL, "engine must be in [0, 6]");
  if (p->H * p->W > (int64_t(1) << 30) - 8)
    return fail(SCIGAP_EINVAL, "H*W must be <= 2^30 - 8 (buffer-addressed planes)");
  return SCIGAP_OK;
}

// Forward of the unrolled model (grad.hip): K accelerated GAP iterations with
// per-iteration lambda_k / weight_k and no early exit (eps = 0 never meets
// |E_prev - E| < eps*E_0).  xs holds x_0 on entry; iteration k reads x_k from
// xs + k*xstride and writes x_{k+1} to xs + (k+1)*xstride (the last one to
// x_out); the residual r_k it uses is written to rs + k*rstride (strides in
// floats).  y1 ping-pongs between the two planes of y1b; nexit holds K ints.
int unrolled_forward_iterations(int64_t H, int64_t W, int64_t T, int K, int D,
                                const float* lambdas, const float* weights, const float* y,
                                const float* phi, const float* S, float* xs, size_t xstride,
                                float* y1b, float* x_out, float* rs, size_t rstride,
                                int* nexit, float* yb_part, hipStream_t s) {
  const int64_t N = H * W;
  if (D < 1 || D > kMaxFused) return fail(SCIGAP_EUNSUPPORTED, "unrolled depth %d", D);
  {
    ProfScope ps("init", s);
    k_init<<<grid_pixels(N), 256, 0, s>>>(xs, phi, y, S, rs, y1b, N, int(T), GAP_ACC, 0.f);
  }
  SCIGAP_LAUNCH_CHECK("unrolled init launch");
  for (int k = 0; k < K; ++k) {
    FusedArgs a{};
    a.x_in = xs + size_t(k) * xstride;
    a.x_out = k + 1 < K ? xs + size_t(k + 1) * xstride : x_out;
    a.phi = phi;
    a.y = y;
    a.phi_sum = S;
    a.r_in = rs + size_t(k) * rstride;
    a.r_out = k + 1 < K ? rs + size_t(k + 1) * rstride : nullptr;
    a.y1_in = y1b + size_t(k & 1) * N;
    a.y1_out = k + 1 < K ? y1b + size_t((k + 1) & 1) * N : nullptr;
    a.eslot = nullptr;  // eps = 0: no early exit, no energies, decision or repair pass
    a.stop = nullptr;
    a.nexit = nexit + k;
    a.yb_part = yb_part;
    a.H = int(H);
    a.W = int(W);
    a.T = int(T);
    a.lam = lambdas[k];
    a.weight = weights[k];
    a.tw = float(0.25 / double(weights[k]));
    a.eps = 0.0;
    a.mode = GAP_ACC;
    a.epilogue = k + 1 < K;
    if (int rc = launch_fused_depth(D, a, s)) return rc;
  }
  return SCIGAP_OK;
}

}  // namespace scigap

using namespace scigap;

extern "C" {

void scigap_default_params(scigap_params* p, int64_t H, int64_t W, int64_t T) {
  if (!p) return;
  p->H = H;
  p->W = W;
  p->T = T;
  p->method = SCIGAP_METHOD_GAP;
  p->accelerate = 1;
  p->clip = 0;
  p->iters = 50;
  p->lambda_ = 1.f;
  p->gamma = 0.01f;
  p->tv_weight = 0.1f;
  p->tv_eps = 2e-4f;
  p->tv_iter_max = 5;
  p->tv_weight_decay = 1.f;
  p->gamma_decay = 1.f;
  p->tile_rows = 0;
  p->engine = 0;
}

size_t scigap_solve_workspace_bytes(const scigap_params* p) {
  if (!p || check_params(p) != SCIGAP_OK) return 0;
  return carve(p, nullptr).bytes;
}

int scigap_solve(const scigap_params* p, const float* y, const float* phi, const float* phi_sum,
                 const float* x0, float* x_out, const float* truth, double* sq_err, int* stats,
                 void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  if (int rc = check_params(p)) return rc;
  if (!y || !phi || !phi_sum || !x_out) return fail(SCIGAP_ENULL, "scigap_solve: null input/output");
  if (truth && !sq_err) return fail(SCIGAP_ENULL, "scigap_solve: truth given without sq_err");
  const int64_t N = p->H * p->W, T = p->T;
  const size_t vol = size_t(N * T) * 4;
  if (x0 && overlaps(x0, vol, x_out, vol)) return fail(SCIGAP_EALIAS, "x0 overlaps x_out");
  if (overlaps(x_out, vol, y, N * 4) || overlaps(x_out, vol, phi, vol) ||
      overlaps(x_out, vol, phi_sum, N * 4))
    return fail(SCIGAP_EALIAS, "x_out overlaps an input");
  Workspace ws = carve(p, static_cast<char*>(workspace));
  if (!workspace || workspace_bytes < ws.bytes)
    return fail(SCIGAP_EWORKSPACE, "workspace needs %zu bytes, got %zu", ws.bytes, workspace_bytes);
  if (overlaps(workspace, ws.bytes, x_out, vol)) return fail(SCIGAP_EALIAS, "workspace overlaps x_out");
  // The tile loads read float2 pairs; keep every plane 8-byte aligned.
  if ((reinterpret_cast<uintptr_t>(x_out) | reinterpret_cast<uintptr_t>(phi) |
       reinterpret_cast<uintptr_t>(y) | reinterpret_cast<uintptr_t>(x0)) % 8 != 0)
    return fail(SCIGAP_EINVAL, "device buffers must be 8-byte aligned");

  const int D = p->tv_iter_max - 1;
  const bool deep = D > kMaxFused;
  const int deep_depth = deep ? tuned_tv_depth(p->H, p->W, p->T, p->tv_iter_max) : 0;
  int tile_rows = p->tile_rows, engine = p->engine;
  if (tile_rows == 0 && engine == 0) tuned_solve_config(*p, &tile_rows, &engine);
  hipStream_t s = as_stream(stream);
  const int K = p->iters;
  const int mode = solver_mode(*p);
  if (K == 0) {
    // No iterations: x = x0 or At(y).
    if (x0) {
      hipError_t e = hipMemcpyAsync(x_out, x0, vol, hipMemcpyDeviceToDevice, s);
      return e == hipSuccess ? SCIGAP_OK : hip_fail(e, "copy x0");
    }
    return scigap_adjoint(y, phi, x_out, p->H, p->W, T, stream);
  }
  if (!deep && D > 0) {
    // early-stop counters of every iteration
    k_fill_i32<<<unsigned((K + 255) / 256), 256, 0, s>>>(ws.st.nexit, K, 0);
    SCIGAP_LAUNCH_CHECK("counter fill");
  }
  if (truth)
    if (int rc = fill_f64(sq_err, K, 0.0, s)) return rc;
  if (stats) {
    k_fill_i32<<<(K + 255) / 256, 256, 0, s>>>(stats, K, 0);
    SCIGAP_LAUNCH_CHECK("stats fill");
  }
  {
    ProfScope ps("init", s);
    k_init<<<grid_pixels(N), 256, 0, s>>>(x0, phi, y, phi_sum, ws.st.r[0], ws.st.y1[0], N, int(T),
                                          mode, float(p->gamma));
  }
  SCIGAP_LAUNCH_CHECK("init launch");

  const int fused_iters = fused_iterations(*p);
  for (int k = 0; k < fused_iters; ++k) {
    FusedArgs a = iteration_args(*p, k, p->H, p->W, T, y, phi, phi_sum, x0, x_out, truth, sq_err,
                                 stats, ws.st);
    if (D == 0) {
      k_project<<<grid_pixels(N), 256, 0, s>>>(a);
      SCIGAP_LAUNCH_CHECK("projection launch");
    } else if (deep) {
      // v -> vbuf, TV(vbuf) -> X_{k+1}, then the residual from (X_{k+1}, vbuf).
      FusedArgs pa = a;
      pa.x_out = ws.vbuf;
      k_gap_proj<<<grid_pixels(N), 256, 0, s>>>(pa);
      SCIGAP_LAUNCH_CHECK("projection launch");
      if (int rc = tv_chunked(ws.vbuf, a.x_out, p->H, p->W, T, a.weight, p->tv_eps,
                              p->tv_iter_max, deep_depth, stats ? ws.st.stop : nullptr,
                              ws.tvws, s))
        return rc;
      if (stats) {  // frames whose TV stopped early in this iteration
        k_count_early<<<1, 256, 0, s>>>(ws.st.stop, int(T), p->tv_iter_max - 1, stats + k);
        SCIGAP_LAUNCH_CHECK("early-stop count launch");
      }
      FusedArgs ea = a;
      ea.x_in = ws.vbuf;
      k_gap_epi<<<grid_pixels(N), 256, 0, s>>>(ea);
      SCIGAP_LAUNCH_CHECK("epilogue launch");
    } else if (int rc = launch_fused_depth(D, a, s, tile_rows, engine)) {
      return rc;
    }
  }
  if (mode == ADMM) {
    // x_{K-1} = theta_{K-1} + b_{K-1} + lambda*r_{K-1}*Phi.
    const bool fy = (K == 1 && !x0);
    const float* th = K == 1 ? x0 : state_buffer(*p, ws.st, x_out, K - 1);
    k_admm_final<<<grid_pixels(N), 256, 0, s>>>(th, K == 1 ? nullptr : ws.st.bbuf,
                                                ws.st.r[(K - 1) & 1], phi, y, x_out, truth,
                                                truth ? sq_err + (K - 1) : nullptr, N, int(T),
                                                p->lambda_, fy ? 1 : 0);
    SCIGAP_LAUNCH_CHECK("admm final launch");
  }
  return SCIGAP_OK;
}

size_t scigap_tv_workspace_bytes_depth(int64_t H, int64_t W, int64_t T, int n_iter_max,
                                       int depth) {
  if (check_shape(H, W, T) != SCIGAP_OK || n_iter_max < 1 || depth < 0 || depth > kMaxFused)
    return 0;
  const int D = n_iter_max - 1;
  if (D == 0) return 256;
  if (depth == 0 && D <= kMaxFused) depth = D;
  if (depth >= D)
    return align256(size_t(T) * D * size_t(max_slot_tiles(D, int(H), int(W))) * 4) +
           align256(size_t(T) * 4) + 256;
  return tv_chunked_workspace(H, W, T, n_iter_max, D);
}

size_t scigap_tv_workspace_bytes(int64_t H, int64_t W, int64_t T, int n_iter_max) {
  return scigap_tv_workspace_bytes_depth(H, W, T, n_iter_max, 0);
}

int scigap_tv_chambolle_depth(const float* image, float* out, int64_t H, int64_t W, int64_t T,
                              float weight, float eps, int n_iter_max, int depth, int* stop_iter,
                              void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  if (int rc = check_shape(H, W, T)) return rc;
  if (!image || !out) return fail(SCIGAP_ENULL, "scigap_tv_chambolle: null image/out");
  if (n_iter_max < 1) return fail(SCIGAP_EINVAL, "n_iter_max must be >= 1");
  if (!(weight > 0.f) || !std::isfinite(weight))
    return fail(SCIGAP_EINVAL, "weight must be finite and > 0");
  if (!std::isfinite(eps)) return fail(SCIGAP_EINVAL, "eps must be finite");
  if (depth < 0 || depth > kMaxFused) return fail(SCIGAP_EINVAL, "depth must be in [0, 6]");
  if (H * W > (int64_t(1) << 30) - 8)
    return fail(SCIGAP_EINVAL, "H*W must be <= 2^30 - 8 (buffer-addressed planes)");
  const size_t vol = size_t(H * W * T) * 4;
  if (overlaps(image, vol, out, vol)) return fail(SCIGAP_EALIAS, "image and out overlap");
  const size_t need = scigap_tv_workspace_bytes_depth(H, W, T, n_iter_max, depth);
  if (!workspace || workspace_bytes < need)
    return fail(SCIGAP_EWORKSPACE, "workspace needs %zu bytes, got %zu", need, workspace_bytes);
  if ((reinterpret_cast<uintptr_t>(image) | reinterpret_cast<uintptr_t>(out)) % 8 != 0)
    return fail(SCIGAP_EINVAL, "device buffers must be 8-byte aligned");
  hipStream_t s = as_stream(stream);
  const int D = n_iter_max - 1;
  if (D == 0) {
    // One iteration returns the input (no dual update happens).
    hipError_t e = hipMemcpyAsync(out, image, vol, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return hip_fail(e, "copy");
    if (stop_iter) {
      k_fill_i32<<<unsigned((T + 255) / 256), 256, 0, s>>>(stop_iter, int(T), 0);
      SCIGAP_LAUNCH_CHECK("stop fill");
    }
    return SCIGAP_OK;
  }
  if (depth == 0) depth = D <= kMaxFused ? D : tuned_tv_depth(H, W, T, n_iter_max);
  if (depth < D)
    return tv_chunked(image, out, H, W, T, weight, eps, n_iter_max, depth, stop_iter, workspace,
                      s);
  // All updates in one launch from p = 0.
  const int stride = max_slot_tiles(D, int(H), int(W));
  float* slots = static_cast<float*>(workspace);
  char* rest = static_cast<char*>(workspace) + align256(size_t(T) * D * size_t(stride) * 4);
  FusedArgs a{};
  a.x_in = image;
  a.x_out = out;
  a.eslot = slots;
  a.eslot_stride = stride;
  a.stop = stop_iter ? stop_iter : reinterpret_cast<int*>(rest);
  a.nexit = reinterpret_cast<int*>(rest + align256(size_t(T) * 4));
  k_fill_i32<<<1, 64, 0, s>>>(a.nexit, 1, 0);
  SCIGAP_LAUNCH_CHECK("counter fill");
  a.H = int(H);
  a.W = int(W);
  a.T = int(T);
  a.weight = weight;
  a.tw = float(0.25 / double(weight));
  a.eps = eps;
  a.mode = TV_ONLY;
  return launch_fused_depth(D, a, s);
}

int scigap_tv_chambolle(const float* image, float* out, int64_t H, int64_t W, int64_t T,
                        float weight, float eps, int n_iter_max, int* stop_iter, void* workspace,
                        size_t workspace_bytes, void* stream) {
  return scigap_tv_chambolle_depth(image, out, H, W, T, weight, eps, n_iter_max, 0, stop_iter,
                                   workspace, workspace_bytes, stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// TV without a caller workspace (scigap_tv_chambolle_noalloc)
// ---------------------------------------------------------------------------

namespace scigap {
namespace {

// The scratch of scigap_tv_chambolle_noalloc: the per-tile energy slots, the
// stop array (when the caller passes none) and the early-stop counter of the
// single-launch TV.  Every GPU the library runs on holds one copy of it until
// the process ends.
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device.  The array's name is a
// host-side stand-in; only this address may be handed to kernels.
int scratch_address(void** ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(SCIGAP_EUNSUPPORTED, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(SCIGAP_EHIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return SCIGAP_OK;
}

}  // namespace
}  // namespace scigap

extern "C" {

size_t scigap_tv_noalloc_max_bytes(void) { return scigap::kNoAllocScratchBytes; }

int scigap_tv_chambolle_noalloc(const float* image, float* out, int64_t H, int64_t W, int64_t T,
                                float weight, float eps, int n_iter_max, int* stop_iter,
                                void* stream) {
  clear_error();
  if (int rc = check_shape(H, W, T)) return rc;
  if (n_iter_max < 1) return fail(SCIGAP_EINVAL, "n_iter_max must be >= 1");
  if (n_iter_max - 1 > kMaxFused)
    return fail(SCIGAP_EUNSUPPORTED,
                "scigap_tv_chambolle_noalloc: n_iter_max %d needs the chunked TV and its "
                "dual-field workspace; at most %d iterations run without one",
                n_iter_max, kMaxFused + 1);
  const size_t need = scigap_tv_workspace_bytes_depth(H, W, T, n_iter_max, 0);
  if (need == 0) return fail(SCIGAP_EINVAL, "scigap_tv_chambolle_noalloc: bad shape");
  if (need > kNoAllocScratchBytes)
    return fail(SCIGAP_EINVAL,
                "scigap_tv_chambolle_noalloc: %lldx%lldx%lld with n_iter_max %d needs %zu bytes "
                "of scratch, more than the %zu the library holds; use scigap_tv_chambolle",
                (long long)H, (long long)W, (long long)T, n_iter_max, need,
                kNoAllocScratchBytes);
  void* ws = nullptr;
  if (int rc = scratch_address(&ws)) return rc;
  return scigap_tv_chambolle_depth(image, out, H, W, T, weight, eps, n_iter_max, 0, stop_iter, ws,
                                   kNoAllocScratchBytes, stream);
}

}  // extern "C"
