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
// Early exit.  Each launch accumulates the per-frame energies E_0..E_{D-1}.
// A second launch of the same kernel with FIX=true reads them; when no frame
// stopped early every workgroup returns at once, otherwise it recomputes the
// frames that stopped at iteration s < D from the unchanged inputs (x, r are
// double-buffered), rewrites them, and recomputes the epilogue for its tile.

#include "tv_tile.h"

namespace scigap {

int fill_f64(double* p, int n, double v, hipStream_t s);
size_t tv_general_workspace(int64_t H, int64_t W, int64_t T, int n);
int tv_general(const float* image, float* out, int64_t H, int64_t W, int64_t T, float weight,
               float eps, int n, int* stop_iter, void* ws, hipStream_t s);
int tv_stop_from_energy(const double* energy, int T, int D, double eps, int* stop, hipStream_t s);

namespace {

// TV_ONLY runs the stencil alone (denoise_tv_chambolle): img = x_in, no
// projection, no residual epilogue.
enum Mode { GAP_ACC = 0, GAP_PLAIN = 1, ADMM = 2, TV_ONLY = 3 };

struct FusedArgs {
  const float* x_in;    // X_k (T,H,W); unused when from_y (X_0 = At(y))
  float* x_out;         // X_{k+1}
  const float* phi;     // (T,H,W)
  const float* y;       // (H,W)
  const float* phi_sum; // (H,W)
  const float* r_in;    // residual of this iteration
  float* r_out;         // residual of the next iteration
  const float* y1_in;   // accelerated GAP: y1 of this iteration (after its update)
  float* y1_out;
  const float* truth;   // optional (T,H,W)
  double* err;          // sum of squared errors of this iteration
  float* b;             // ADMM b (T,H,W)
  double* energy;       // [T][D] energies of this iteration
  int* stats;           // frames that stopped early (FIX adds)
  int H, W, T;
  float lam, gam_next, weight, tw;
  double eps;
  int mode, clip, from_y, write_b, read_b, epilogue;
  int tiles_x, ntiles, per_xcd;
};

// XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs, so
// give XCD i a contiguous band of tiles and let neighbouring tiles (which
// re-read each other's halos) share that XCD's L2.
__device__ __forceinline__ int tile_of_block(int bid, int per_xcd) {
  return (bid & 7) * per_xcd + (bid >> 3);
}

// Load an RY x 2 register block of one plane (zero outside the image).
// `off` is the plane offset of register (0,0).  INSIDE tiles lie entirely in
// the image (and W is even), so they use unguarded 8-byte loads with unsigned
// offsets (scalar base + 32-bit vector offset addressing).
template <int RY, bool INSIDE>
__device__ __forceinline__ void load_plane(const float* __restrict__ src, float (&v)[RY][2],
                                           const LanePos& lp, int off, int H, int W) {
  if constexpr (INSIDE) {
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      float2 q = *reinterpret_cast<const float2*>(src + unsigned(off + j * W));
      v[j][0] = q.x;
      v[j][1] = q.y;
    }
  } else {
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      int r = lp.row0 + j;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        int c = lp.col + k;
        v[j][k] = (r >= 0 && r < H && c >= 0 && c < W) ? src[unsigned(off + j * W + k)] : 0.f;
      }
    }
  }
}

// Tile placement shared by the main and repair passes.
template <int D, int BY, int MR>
struct TilePlace {
  using G = TileGeom<D, BY, MR>;
  LanePos lp;
  int off0;        // plane offset of the lane's register (0,0)
  bool inside;     // load tile entirely inside the image and W even
  bool lane_owner; // lane's columns are interior columns of the tile
  bool own_c[2];   // interior column inside the image
  unsigned own_r;  // bit j: row0+j is an interior row inside the image
  __device__ TilePlace(int tile, int tiles_x, int H, int W) {
    const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
    const int c0 = tx * G::OX - G::HC, r0 = ty * G::OY - D;
    lp.w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    lp.lane = threadIdx.x & 63;
    lp.row0 = r0 + lp.w * G::RY;
    lp.col = c0 + 2 * lp.lane;
    off0 = lp.row0 * W + lp.col;
    inside = c0 >= 0 && r0 >= 0 && c0 + G::LW <= W && r0 + G::LH <= H && (W % 2 == 0);
    lane_owner = lp.lane >= G::HC / 2 && lp.lane < 64 - G::HC / 2;
    own_c[0] = lane_owner && lp.col < W;
    own_c[1] = lane_owner && lp.col + 1 < W;
    own_r = 0;
#pragma unroll
    for (int j = 0; j < G::RY; ++j) {
      const int tr = lp.w * G::RY + j;  // row within the load tile
      own_r |= unsigned(tr >= D && tr < G::LH - D && lp.row0 + j < H) << j;
    }
  }
  __device__ bool owns(int j, int k) const { return own_c[k] && ((own_r >> j) & 1u); }
};

// Residual of the next iteration at one owned pixel.
template <int MODE>
__device__ __forceinline__ void epilogue_pixel(const FusedArgs& a, unsigned o, float yb) {
  float yv = a.y[o], s = a.phi_sum[o];
  if constexpr (MODE == GAP_ACC) {
    float y1 = a.y1_in[o] + (yv - yb);
    a.y1_out[o] = y1;
    a.r_out[o] = (y1 - yb) / s;
  } else if constexpr (MODE == GAP_PLAIN) {
    a.r_out[o] = (yv - yb) / s;
  } else {
    a.r_out[o] = (yv - yb) / (s + a.gam_next);
  }
}

template <int BY>
__device__ __forceinline__ void commit_err(float err, double* dst, double* red64, int w, int lane) {
  double e = wave_sum(double(err));
  if (lane == 0) red64[w] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int q = 0; q < BY; ++q) tot += red64[q];
    atomicAdd(dst, tot);
  }
}

// Main pass.  INSIDE selects the interior tiles (no masks, no bounds checks)
// or the border ring; IQA adds the squared error against the truth.
// Registers hold the stencil state (img, out, p0, p1) and the prefetched next
// frame (x, Phi); lambda*r, the current Phi and the yb accumulator, each used
// once per frame, live in LDS as [tile row][128] float planes.
template <int D, int BY, int MR, int MODE, bool INSIDE, bool IQA>
__global__ void __launch_bounds__(64 * BY) k_fused(FusedArgs a) {
  using G = TileGeom<D, BY, MR>;
  constexpr int RY = G::RY;
  constexpr int NH = MR;
  __shared__ TileLds<BY> lds;
  __shared__ float2 s_q[G::LH * 64], s_ph[G::LH * 64], s_yb[G::LH * 64], s_img[G::LH * 64];
  __shared__ float red[BY * D];
  __shared__ double red64[BY];

  const int tile = tile_of_block(blockIdx.x, a.per_xcd);
  if (tile >= a.ntiles) return;
  const int H = a.H, W = a.W, T = a.T;
  const int N = H * W;
  const TilePlace<D, BY, MR> tp(tile, a.tiles_x, H, W);
  if (tp.inside != INSIDE) return;
  const LanePos& lp = tp.lp;
  const int off0 = tp.off0;
  // LDS slot of register row j of this lane: (w*RY + j)*64 + lane.
  const int sl0 = lp.w * RY * 64 + lp.lane;

  EdgeMasks em;
  if constexpr (!INSIDE) em.init<RY>(lp, H, W);

  const bool from_y = a.from_y;
  if constexpr (MODE != TV_ONLY) {
    // q = lambda*r (or y + lambda*r when X_0 = At(y)) over the load tile.
    float q[RY][2];
    load_plane<RY, INSIDE>(a.r_in, q, lp, off0, H, W);
    if (from_y) {
      float yv[RY][2];
      load_plane<RY, INSIDE>(a.y, yv, lp, off0, H, W);
#pragma unroll
      for (int j = 0; j < RY; ++j)
        s_q[sl0 + j * 64] = make_float2(fmaf(a.lam, q[j][0], yv[j][0]), fmaf(a.lam, q[j][1], yv[j][1]));
    } else {
#pragma unroll
      for (int j = 0; j < RY; ++j) s_q[sl0 + j * 64] = make_float2(a.lam * q[j][0], a.lam * q[j][1]);
    }
#pragma unroll
    for (int j = 0; j < RY; ++j) s_yb[sl0 + j * 64] = make_float2(0.f, 0.f);
  }
  float err = 0.f;

  // Frame t+1's x and Phi are requested before frame t's TV so that their
  // latency hides behind the stencil arithmetic.
  float nx[RY][2], nph[RY][2];
  const float* phit = a.phi;
  const float* xt = a.x_in;
  if constexpr (MODE != TV_ONLY) load_plane<RY, INSIDE>(phit, nph, lp, off0, H, W);
  if (!from_y) load_plane<RY, INSIDE>(xt, nx, lp, off0, H, W);

  float* xo = a.x_out;
  const float* trt = a.truth;
  float* bt = a.b;
  double* en = a.energy;
  // Every lane reads back only the LDS slots it wrote itself, so the planes
  // s_q, s_ph, s_img and s_yb need no barriers.
  float2* const img = s_img + sl0;
  for (int t = 0; t < T; ++t) {
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      if constexpr (MODE == TV_ONLY) {
        img[j * 64] = make_float2(nx[j][0], nx[j][1]);
      } else {
        const float2 qj = s_q[sl0 + j * 64];
        s_ph[sl0 + j * 64] = make_float2(nph[j][0], nph[j][1]);
        img[j * 64] = make_float2(from_y ? qj.x * nph[j][0] : fmaf(qj.x, nph[j][0], nx[j][0]),
                                  from_y ? qj.y * nph[j][1] : fmaf(qj.y, nph[j][1], nx[j][1]));
      }
    }
    if (t + 1 < T) {
      if constexpr (MODE != TV_ONLY) load_plane<RY, INSIDE>(phit + N, nph, lp, off0, H, W);
      if (!from_y) load_plane<RY, INSIDE>(xt + N, nx, lp, off0, H, W);
    }

    float out[RY][2];
    float acc[NH * 8];
    tv_run_from_zero<D, RY, BY, NH, !INSIDE>(img, out, D, lds, lp, em, a.weight, a.tw, acc, true);

#pragma unroll
    for (int j = 0; j < RY; ++j) {
      if (!((tp.own_r >> j) & 1u)) continue;
      if constexpr (MODE == TV_ONLY) {
#pragma unroll
        for (int k = 0; k < 2; ++k)
          if (tp.own_c[k]) xo[unsigned(off0 + j * W + k)] = out[j][k];
        continue;
      }
      const float2 ph = s_ph[sl0 + j * 64];
      const float2 im = img[j * 64];
      float2 yb = s_yb[sl0 + j * 64];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (!tp.own_c[k]) continue;
        const unsigned o = unsigned(off0 + j * W + k);
        const float f = k ? ph.y : ph.x;
        const float v = k ? im.y : im.x;
        float& ybk = k ? yb.y : yb.x;
        float xn = out[j][k];
        if constexpr (MODE == ADMM) {
          if (a.clip) xn = fminf(fmaxf(xn, 0.f), 1.f);
          if constexpr (IQA) {  // x_k = v_k + b_k
            float bk = a.read_b ? bt[o] : 0.f;
            float e = trt[o] - (v + bk);
            err = fmaf(e, e, err);
          }
          if (a.write_b) bt[o] = xn - v;
          ybk = fmaf(f, 2.f * xn - v, ybk);
        } else {
          if constexpr (IQA) {
            float e = trt[o] - xn;
            err = fmaf(e, e, err);
          }
          ybk = fmaf(f, xn, ybk);
        }
        xo[o] = xn;
      }
      s_yb[sl0 + j * 64] = yb;
    }
    energy_commit<D, BY, MR, NH>(acc, tp.lane_owner, en, red, lp);
    phit += N;
    xt += N;
    xo += N;
    en += D;
    if constexpr (IQA) trt += N;
    if constexpr (MODE == ADMM) bt += (bt ? N : 0);
  }

  if (MODE != TV_ONLY && a.epilogue) {
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      if (!((tp.own_r >> j) & 1u)) continue;
      const float2 yb = s_yb[sl0 + j * 64];
#pragma unroll
      for (int k = 0; k < 2; ++k)
        if (tp.own_c[k]) epilogue_pixel<MODE>(a, unsigned(off0 + j * W + k), k ? yb.y : yb.x);
    }
  }
  if constexpr (IQA) commit_err<BY>(err, a.err, red64, lp.w, lp.lane);
}

// Early-exit repair pass (rarely does work).  Every workgroup first derives
// the exit iteration of every frame from the energies of the main pass; if no
// frame stopped early it returns.  Otherwise it recomputes the frames that
// stopped at s < D from the unchanged inputs, rewrites them, and recomputes
// the epilogue of its tile from all frames.
template <int D, int BY, int MR>
__global__ void __launch_bounds__(64 * BY) k_fused_fix(FusedArgs a) {
  using G = TileGeom<D, BY, MR>;
  constexpr int RY = G::RY;
  constexpr int NH = MR;
  __shared__ TileLds<BY> lds;
  __shared__ float2 s_img[G::LH * 64];
  __shared__ double red64[BY];

  const int tile = tile_of_block(blockIdx.x, a.per_xcd);
  if (tile >= a.ntiles) return;
  const int H = a.H, W = a.W, T = a.T;
  const int N = H * W;
  int nexit = 0;
  for (int t = 0; t < T; ++t) nexit += exit_iteration<D>(a.energy + size_t(t) * D, a.eps) < D;
  if (nexit == 0) return;
  if (tile == 0 && threadIdx.x == 0 && a.stats) atomicAdd(a.stats, nexit);

  const TilePlace<D, BY, MR> tp(tile, a.tiles_x, H, W);
  const LanePos& lp = tp.lp;
  const int off0 = tp.off0;
  EdgeMasks em;
  em.init<RY>(lp, H, W);

  const bool tv_only = a.mode == TV_ONLY;
  float q[RY][2];
  if (tv_only) {
#pragma unroll
    for (int j = 0; j < RY; ++j) q[j][0] = q[j][1] = 0.f;
  } else {
    load_plane<RY, false>(a.r_in, q, lp, off0, H, W);
  }
  if (a.from_y) {
    float yv[RY][2];
    load_plane<RY, false>(a.y, yv, lp, off0, H, W);
#pragma unroll
    for (int j = 0; j < RY; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) q[j][k] = fmaf(a.lam, q[j][k], yv[j][k]);
  } else {
#pragma unroll
    for (int j = 0; j < RY; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) q[j][k] *= a.lam;
  }
  float yb[RY][2];
#pragma unroll
  for (int j = 0; j < RY; ++j) yb[j][0] = yb[j][1] = 0.f;
  float err = 0.f;

  for (int t = 0; t < T; ++t) {
    const float* phit = a.phi + size_t(t) * N;
    const float* xt = a.x_in + size_t(t) * N;
    float* xo = a.x_out + size_t(t) * N;
    const int nup = exit_iteration<D>(a.energy + size_t(t) * D, a.eps);
    if (nup == D && tv_only) continue;
    if (nup == D) {
      // Frame was right: only its contribution to the epilogue is needed.
#pragma unroll
      for (int j = 0; j < RY; ++j)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          if (!tp.owns(j, k)) continue;
          const unsigned o = unsigned(off0 + j * W + k);
          float f = phit[o], xn = xo[o];
          if (a.mode == ADMM) {
            float v = a.from_y ? q[j][k] * f : fmaf(q[j][k], f, xt[o]);
            yb[j][k] = fmaf(f, 2.f * xn - v, yb[j][k]);
          } else {
            yb[j][k] = fmaf(f, xn, yb[j][k]);
          }
        }
      continue;
    }
    float ph[RY][2], img[RY][2], xv[RY][2];
    if (tv_only) {
#pragma unroll
      for (int j = 0; j < RY; ++j) ph[j][0] = ph[j][1] = 0.f;
    } else {
      load_plane<RY, false>(phit, ph, lp, off0, H, W);
    }
    if (!a.from_y) load_plane<RY, false>(xt, xv, lp, off0, H, W);
    float2* const simg = s_img + lp.w * RY * 64 + lp.lane;
#pragma unroll
    for (int j = 0; j < RY; ++j) {
#pragma unroll
      for (int k = 0; k < 2; ++k)
        img[j][k] = a.from_y ? q[j][k] * ph[j][k] : fmaf(q[j][k], ph[j][k], xv[j][k]);
      simg[j * 64] = make_float2(img[j][0], img[j][1]);
    }
    float out[RY][2];
    float acc[NH * 8];
    tv_run_from_zero<D, RY, BY, NH, true>(simg, out, nup, lds, lp, em, a.weight, a.tw, acc, false);
    const float* trt = a.truth ? a.truth + size_t(t) * N : nullptr;
    float* bt = a.b ? a.b + size_t(t) * N : nullptr;
#pragma unroll
    for (int j = 0; j < RY; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (!tp.owns(j, k)) continue;
        const unsigned o = unsigned(off0 + j * W + k);
        float xn = out[j][k];
        if (tv_only) {
          xo[o] = xn;
          continue;
        }
        if (a.mode == ADMM) {
          if (a.clip) xn = fminf(fmaxf(xn, 0.f), 1.f);
          if (a.write_b) bt[o] = xn - img[j][k];
          yb[j][k] = fmaf(ph[j][k], 2.f * xn - img[j][k], yb[j][k]);
        } else {
          if (trt) {  // replace the main pass's error of this pixel
            float e = trt[o] - xn, eo = trt[o] - xo[o];
            err = fmaf(e, e, err);
            err = fmaf(-eo, eo, err);
          }
          yb[j][k] = fmaf(ph[j][k], xn, yb[j][k]);
        }
        xo[o] = xn;
      }
  }
  if (!tv_only && a.epilogue) {
#pragma unroll
    for (int j = 0; j < RY; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (!tp.owns(j, k)) continue;
        const unsigned o = unsigned(off0 + j * W + k);
        if (a.mode == GAP_ACC) epilogue_pixel<GAP_ACC>(a, o, yb[j][k]);
        else if (a.mode == GAP_PLAIN) epilogue_pixel<GAP_PLAIN>(a, o, yb[j][k]);
        else epilogue_pixel<ADMM>(a, o, yb[j][k]);
      }
  }
  if (a.truth && a.err && a.mode != ADMM && !tv_only) commit_err<BY>(err, a.err, red64, lp.w, lp.lane);
}

// Initial residual: X_0 = x0 (or At(y)), yb = A(X_0), then
// GAP-acc: y1 = y - yb, r = (y1 - yb)/Phi_sum; GAP-plain: r = (y - yb)/Phi_sum;
// ADMM: r = (y - yb)/(Phi_sum + gamma)   (pnp_sci_algo.py:625-633, 640-645, 808-809)
__global__ void __launch_bounds__(256) k_init(const float* __restrict__ x0,
                                              const float* __restrict__ phi,
                                              const float* __restrict__ y,
                                              const float* __restrict__ phi_sum,
                                              float* __restrict__ r, float* __restrict__ y1,
                                              int64_t N, int T, int mode, float gam) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < N;
       i += int64_t(gridDim.x) * blockDim.x) {
    float yv = y[i], yb = 0.f;
    for (int t = 0; t < T; ++t) {
      float f = phi[t * N + i];
      float xv = x0 ? x0[t * N + i] : yv * f;
      yb = fmaf(xv, f, yb);
    }
    float s = phi_sum[i];
    if (mode == GAP_ACC) {
      float v1 = yv - yb;
      y1[i] = v1;
      r[i] = (v1 - yb) / s;
    } else if (mode == GAP_PLAIN) {
      r[i] = (yv - yb) / s;
    } else {
      r[i] = (yv - yb) / (s + gam);
    }
  }
}

// Last ADMM iteration: x = theta + b + lambda*r*Phi (pnp_sci_algo.py:809), in
// place on x_io when theta lives there, and its squared error.
__global__ void __launch_bounds__(256) k_admm_final(const float* __restrict__ theta,
                                                    const float* __restrict__ b,
                                                    const float* __restrict__ r,
                                                    const float* __restrict__ phi,
                                                    const float* __restrict__ y, float* x_out,
                                                    const float* __restrict__ truth, double* err,
                                                    int64_t N, int T, float lam, int from_y) {
  float e = 0.f;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < N;
       i += int64_t(gridDim.x) * blockDim.x) {
    float lr = lam * r[i];
    float yv = from_y ? y[i] : 0.f;
    for (int t = 0; t < T; ++t) {
      size_t o = size_t(t) * N + i;
      float f = phi[o];
      float th = from_y ? yv * f : theta[o];
      float x = fmaf(lr, f, th + (b ? b[o] : 0.f));
      x_out[o] = x;
      if (truth) {
        float d = truth[o] - x;
        e = fmaf(d, d, e);
      }
    }
  }
  if (truth && err) {
    double s = wave_sum(double(e));
    if ((threadIdx.x & 63) == 0) atomicAdd(err, s);
  }
}

// Outer iteration with an identity denoiser (tv_iter_max = 1: scikit-image
// returns its input unchanged): x' = x + lambda*r*Phi per pixel, then the
// next residual.  Pointwise, so x may be updated in place.
__global__ void __launch_bounds__(256) k_project(FusedArgs a) {
  const int64_t N = int64_t(a.H) * a.W;
  float err = 0.f;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < N;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float lr = a.lam * a.r_in[i];
    const float yv = a.y[i];
    float yb = 0.f;
    for (int t = 0; t < a.T; ++t) {
      const size_t o = size_t(t) * N + i;
      const float f = a.phi[o];
      const float xk = a.from_y ? yv * f : a.x_in[o];
      const float v = fmaf(lr, f, xk);
      float xn = v;
      if (a.mode == ADMM) {
        if (a.clip) xn = fminf(fmaxf(xn, 0.f), 1.f);
        if (a.truth) {
          float e = a.truth[o] - (v + (a.read_b ? a.b[o] : 0.f));
          err = fmaf(e, e, err);
        }
        if (a.write_b) a.b[o] = xn - v;
        yb = fmaf(f, 2.f * xn - v, yb);
      } else {
        if (a.truth) {
          float e = a.truth[o] - xn;
          err = fmaf(e, e, err);
        }
        yb = fmaf(f, xn, yb);
      }
      a.x_out[o] = xn;
    }
    if (a.epilogue) {
      if (a.mode == GAP_ACC) epilogue_pixel<GAP_ACC>(a, unsigned(i), yb);
      else if (a.mode == GAP_PLAIN) epilogue_pixel<GAP_PLAIN>(a, unsigned(i), yb);
      else epilogue_pixel<ADMM>(a, unsigned(i), yb);
    }
  }
  if (a.truth && a.err) {
    double s = wave_sum(double(err));
    if ((threadIdx.x & 63) == 0) atomicAdd(a.err, s);
  }
}

// Deep-TV path (tv_iter_max-1 beyond the fused depth), step 1: the TV input
// v = x + lambda*r*Phi (GAP) or theta + lambda*r*Phi (ADMM), written to x_out.
__global__ void __launch_bounds__(256) k_gap_proj(FusedArgs a) {
  const int64_t N = int64_t(a.H) * a.W;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < N;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float lr = a.lam * a.r_in[i];
    const float yv = a.from_y ? a.y[i] : 0.f;
    for (int t = 0; t < a.T; ++t) {
      const size_t o = size_t(t) * N + i;
      const float f = a.phi[o];
      a.x_out[o] = fmaf(lr, f, a.from_y ? yv * f : a.x_in[o]);
    }
  }
}

// Deep-TV path, step 3: with x_out = TV(v) and v in `vin` (passed as x_in),
// clip/b/error bookkeeping and the next residual, as in the fused epilogue.
__global__ void __launch_bounds__(256) k_gap_epi(FusedArgs a) {
  const int64_t N = int64_t(a.H) * a.W;
  const float* vin = a.x_in;
  float err = 0.f;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < N;
       i += int64_t(gridDim.x) * blockDim.x) {
    float yb = 0.f;
    for (int t = 0; t < a.T; ++t) {
      const size_t o = size_t(t) * N + i;
      const float f = a.phi[o];
      float xn = a.x_out[o];
      if (a.mode == ADMM) {
        const float v = vin[o];
        if (a.clip) {
          xn = fminf(fmaxf(xn, 0.f), 1.f);
          a.x_out[o] = xn;
        }
        if (a.truth) {
          float e = a.truth[o] - (v + (a.read_b ? a.b[o] : 0.f));
          err = fmaf(e, e, err);
        }
        if (a.write_b) a.b[o] = xn - v;
        yb = fmaf(f, 2.f * xn - v, yb);
      } else {
        if (a.truth) {
          float e = a.truth[o] - xn;
          err = fmaf(e, e, err);
        }
        yb = fmaf(f, xn, yb);
      }
    }
    if (a.epilogue) {
      if (a.mode == GAP_ACC) epilogue_pixel<GAP_ACC>(a, unsigned(i), yb);
      else if (a.mode == GAP_PLAIN) epilogue_pixel<GAP_PLAIN>(a, unsigned(i), yb);
      else epilogue_pixel<ADMM>(a, unsigned(i), yb);
    }
  }
  if (a.truth && a.err) {
    double s = wave_sum(double(err));
    if ((threadIdx.x & 63) == 0) atomicAdd(a.err, s);
  }
}

__global__ void k_fill_i32(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

constexpr int kMaxFused = 6;

// Workgroup shape per fused depth: 8 waves (512 threads, so the stencil may
// use up to 256 VGPRs per lane); two strips of D rows per lane while D <= 4,
// one strip beyond.  Load tile 128 x 8*MR*D, interior (128-2*HC) x (8*MR*D-2D).
template <int D> struct FusedShape {
  static constexpr int BY = D <= 4 ? 16 : 8;
  static constexpr int MR = 1;
};

template <int D, int MODE, bool IQA>
int launch_main(const FusedArgs& a, unsigned grid, hipStream_t s) {
  constexpr int BY = FusedShape<D>::BY, MR = FusedShape<D>::MR;
  {
    ProfScope ps(MODE == TV_ONLY ? "tv_fused_interior" : "fused_interior", s);
    k_fused<D, BY, MR, MODE, true, IQA><<<grid, 64 * BY, 0, s>>>(a);
  }
  SCIGAP_LAUNCH_CHECK("fused iteration launch (interior tiles)");
  {
    ProfScope ps(MODE == TV_ONLY ? "tv_fused_border" : "fused_border", s);
    k_fused<D, BY, MR, MODE, false, IQA><<<grid, 64 * BY, 0, s>>>(a);
  }
  SCIGAP_LAUNCH_CHECK("fused iteration launch (border tiles)");
  return SCIGAP_OK;
}

template <int D>
int launch_fused(const FusedArgs& a0, hipStream_t s) {
  constexpr int BY = FusedShape<D>::BY, MR = FusedShape<D>::MR;
  using G = TileGeom<D, BY, MR>;
  FusedArgs a = a0;
  a.tiles_x = (a.W + G::OX - 1) / G::OX;
  int tiles_y = (a.H + G::OY - 1) / G::OY;
  a.ntiles = a.tiles_x * tiles_y;
  a.per_xcd = (a.ntiles + 7) / 8;
  unsigned grid = unsigned(a.per_xcd * 8);
  const bool iqa = a.truth != nullptr;
  int rc;
  if (a.mode == GAP_ACC)
    rc = iqa ? launch_main<D, GAP_ACC, true>(a, grid, s) : launch_main<D, GAP_ACC, false>(a, grid, s);
  else if (a.mode == GAP_PLAIN)
    rc = iqa ? launch_main<D, GAP_PLAIN, true>(a, grid, s) : launch_main<D, GAP_PLAIN, false>(a, grid, s);
  else if (a.mode == ADMM)
    rc = iqa ? launch_main<D, ADMM, true>(a, grid, s) : launch_main<D, ADMM, false>(a, grid, s);
  else
    rc = launch_main<D, TV_ONLY, false>(a, grid, s);
  if (rc) return rc;
  {
    ProfScope ps("fused_fix", s);
    k_fused_fix<D, BY, MR><<<grid, 64 * BY, 0, s>>>(a);
  }
  SCIGAP_LAUNCH_CHECK("fused early-exit launch");
  return SCIGAP_OK;
}

int launch_fused_depth(int D, const FusedArgs& a, hipStream_t s) {
  switch (D) {
    case 1: return launch_fused<1>(a, s);
    case 2: return launch_fused<2>(a, s);
    case 3: return launch_fused<3>(a, s);
    case 4: return launch_fused<4>(a, s);
    case 5: return launch_fused<5>(a, s);
    case 6: return launch_fused<6>(a, s);
  }
  return fail(SCIGAP_EUNSUPPORTED, "fused depth %d not instantiated", D);
}

unsigned grid_pixels(int64_t N) {
  int64_t b = (N + 255) / 256;
  return unsigned(b < 4096 ? (b < 1 ? 1 : b) : 4096);
}

struct Workspace {
  float* xbuf = nullptr;  // T*N
  float* bbuf = nullptr;  // T*N (ADMM)
  float* vbuf = nullptr;  // T*N TV input (deep-TV path)
  void* tvws = nullptr;   // general TV workspace (deep-TV path)
  float* r[2] = {nullptr, nullptr};
  float* y1[2] = {nullptr, nullptr};
  double* energy = nullptr;  // iters * T * D
  size_t bytes = 0;
};

Workspace carve(const scigap_params* p, char* base) {
  Workspace w;
  const size_t N = size_t(p->H) * p->W, T = size_t(p->T);
  const int D = p->tv_iter_max - 1 > 0 ? p->tv_iter_max - 1 : 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* q = base ? base + off : nullptr;
    off += align256(bytes);
    return q;
  };
  w.xbuf = reinterpret_cast<float*>(take(T * N * 4));
  if (p->method == SCIGAP_METHOD_ADMM) w.bbuf = reinterpret_cast<float*>(take(T * N * 4));
  for (int i = 0; i < 2; ++i) w.r[i] = reinterpret_cast<float*>(take(N * 4));
  if (p->method == SCIGAP_METHOD_GAP && p->accelerate)
    for (int i = 0; i < 2; ++i) w.y1[i] = reinterpret_cast<float*>(take(N * 4));
  w.energy = reinterpret_cast<double*>(take(size_t(p->iters > 0 ? p->iters : 1) * T * D * 8));
  if (p->tv_iter_max - 1 > kMaxFused) {
    w.vbuf = reinterpret_cast<float*>(take(T * N * 4));
    w.tvws = take(tv_general_workspace(p->H, p->W, p->T, p->tv_iter_max));
  }
  w.bytes = off;
  return w;
}

int check_params(const scigap_params* p) {
  if (!p) return fail(SCIGAP_ENULL, "params is NULL");
  if (int rc = check_shape(p->H, p->W, p->T)) return rc;
  if (p->method != SCIGAP_METHOD_GAP && p->method != SCIGAP_METHOD_ADMM)
    return fail(SCIGAP_EINVAL, "unknown method %d", p->method);
  if (p->iters < 0) return fail(SCIGAP_EINVAL, "iters must be >= 0");
  if (p->tv_iter_max < 1) return fail(SCIGAP_EINVAL, "tv_iter_max must be >= 1");
  if (!(p->tv_weight > 0.f)) return fail(SCIGAP_EINVAL, "tv_weight must be > 0");
  return SCIGAP_OK;
}

}  // namespace

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

  hipStream_t s = as_stream(stream);
  const int K = p->iters;
  const int mode = p->method == SCIGAP_METHOD_ADMM ? ADMM : (p->accelerate ? GAP_ACC : GAP_PLAIN);
  if (K == 0) {
    // No iterations: x = x0 or At(y).
    if (x0) {
      hipError_t e = hipMemcpyAsync(x_out, x0, vol, hipMemcpyDeviceToDevice, s);
      return e == hipSuccess ? SCIGAP_OK : hip_fail(e, "copy x0");
    }
    return scigap_adjoint(y, phi, x_out, p->H, p->W, T, stream);
  }
  if (!deep)
    if (int rc = fill_f64(ws.energy, K * int(T) * D, 0.0, s)) return rc;
  if (truth)
    if (int rc = fill_f64(sq_err, K, 0.0, s)) return rc;
  if (stats) {
    k_fill_i32<<<(K + 255) / 256, 256, 0, s>>>(stats, K, 0);
    SCIGAP_LAUNCH_CHECK("stats fill");
  }

  // gamma/tv_weight schedule (decays as in ADMM_TV_rec, pnp_sci_algo.py:898-899).
  auto gamma_at = [&](int k) {
    double g = p->gamma;
    for (int i = 0; i < k; ++i) g *= double(p->gamma_decay);
    return g;
  };
  auto weight_at = [&](int k) {
    double w = p->tv_weight;
    for (int i = 0; i < k; ++i) w *= double(p->tv_weight_decay);
    return w;
  };

  {
    ProfScope ps("init", s);
    k_init<<<grid_pixels(N), 256, 0, s>>>(x0, phi, y, phi_sum, ws.r[0], ws.y1[0], N, int(T), mode,
                                          float(gamma_at(0)));
  }
  SCIGAP_LAUNCH_CHECK("init launch");

  const bool iqa = truth != nullptr;
  // ADMM's TV of the last iteration never reaches the returned x.
  const int fused_iters = mode == ADMM ? K - 1 : K;
  // Buffer holding X_j (GAP: x, ADMM: theta): X_last lands in x_out.
  const int last = fused_iters;  // index of the final state
  auto state = [&](int j) -> float* { return ((last - j) % 2 == 0) ? x_out : ws.xbuf; };

  for (int k = 0; k < fused_iters; ++k) {
    FusedArgs a{};
    a.from_y = (k == 0 && !x0);
    a.x_in = k == 0 ? x0 : state(k);
    a.x_out = state(k + 1);
    a.phi = phi;
    a.y = y;
    a.phi_sum = phi_sum;
    a.r_in = ws.r[k & 1];
    a.r_out = ws.r[(k + 1) & 1];
    a.y1_in = ws.y1[k & 1];
    a.y1_out = ws.y1[(k + 1) & 1];
    a.truth = truth;
    a.err = iqa ? sq_err + k : nullptr;
    a.b = ws.bbuf;
    a.energy = ws.energy + size_t(k) * T * D;
    a.stats = stats ? stats + k : nullptr;
    a.H = int(p->H);
    a.W = int(p->W);
    a.T = int(T);
    a.lam = p->lambda_;
    a.gam_next = float(gamma_at(k + 1));
    double wk = weight_at(k);
    a.weight = float(wk);
    a.tw = float(0.25 / wk);
    a.eps = p->tv_eps;
    a.mode = mode;
    a.clip = p->clip;
    a.write_b = mode == ADMM && (iqa || k == fused_iters - 1);
    a.read_b = mode == ADMM && iqa && k > 0;
    a.epilogue = (k + 1 < K);
    if (D == 0) {
      k_project<<<grid_pixels(N), 256, 0, s>>>(a);
      SCIGAP_LAUNCH_CHECK("projection launch");
    } else if (deep) {
      // v -> vbuf, TV(vbuf) -> X_{k+1}, then the residual from (X_{k+1}, vbuf).
      FusedArgs pa = a;
      pa.x_out = ws.vbuf;
      k_gap_proj<<<grid_pixels(N), 256, 0, s>>>(pa);
      SCIGAP_LAUNCH_CHECK("projection launch");
      if (int rc = tv_general(ws.vbuf, a.x_out, p->H, p->W, T, a.weight, p->tv_eps,
                              p->tv_iter_max, nullptr, ws.tvws, s))
        return rc;
      FusedArgs ea = a;
      ea.x_in = ws.vbuf;
      k_gap_epi<<<grid_pixels(N), 256, 0, s>>>(ea);
      SCIGAP_LAUNCH_CHECK("epilogue launch");
    } else if (int rc = launch_fused_depth(D, a, s)) {
      return rc;
    }
  }
  if (mode == ADMM) {
    // x_{K-1} = theta_{K-1} + b_{K-1} + lambda*r_{K-1}*Phi.
    const bool fy = (K == 1 && !x0);
    const float* th = K == 1 ? x0 : state(K - 1);
    k_admm_final<<<grid_pixels(N), 256, 0, s>>>(th, K == 1 ? nullptr : ws.bbuf, ws.r[(K - 1) & 1],
                                                phi, y, x_out, truth, iqa ? sq_err + (K - 1) : nullptr,
                                                N, int(T), p->lambda_, fy ? 1 : 0);
    SCIGAP_LAUNCH_CHECK("admm final launch");
  }
  return SCIGAP_OK;
}

size_t scigap_tv_workspace_bytes(int64_t H, int64_t W, int64_t T, int n_iter_max) {
  if (check_shape(H, W, T) != SCIGAP_OK || n_iter_max < 1) return 0;
  const int D = n_iter_max - 1;
  if (D == 0) return 256;
  if (D <= kMaxFused) return align256(size_t(T) * D * 8);
  return tv_general_workspace(H, W, T, n_iter_max);
}

int scigap_tv_chambolle(const float* image, float* out, int64_t H, int64_t W, int64_t T,
                        float weight, float eps, int n_iter_max, int* stop_iter, void* workspace,
                        size_t workspace_bytes, void* stream) {
  clear_error();
  if (int rc = check_shape(H, W, T)) return rc;
  if (!image || !out) return fail(SCIGAP_ENULL, "scigap_tv_chambolle: null image/out");
  if (n_iter_max < 1) return fail(SCIGAP_EINVAL, "n_iter_max must be >= 1");
  if (!(weight > 0.f)) return fail(SCIGAP_EINVAL, "weight must be > 0");
  const size_t vol = size_t(H * W * T) * 4;
  if (overlaps(image, vol, out, vol)) return fail(SCIGAP_EALIAS, "image and out overlap");
  const size_t need = scigap_tv_workspace_bytes(H, W, T, n_iter_max);
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
  if (D > kMaxFused)
    return tv_general(image, out, H, W, T, weight, eps, n_iter_max, stop_iter, workspace, s);
  double* energy = static_cast<double*>(workspace);
  if (int rc = fill_f64(energy, int(T) * D, 0.0, s)) return rc;
  FusedArgs a{};
  a.x_in = image;
  a.x_out = out;
  a.energy = energy;
  a.H = int(H);
  a.W = int(W);
  a.T = int(T);
  a.weight = weight;
  a.tw = float(0.25 / double(weight));
  a.eps = eps;
  a.mode = TV_ONLY;
  if (int rc = launch_fused_depth(D, a, s)) return rc;
  if (stop_iter) return tv_stop_from_energy(energy, int(T), D, eps, stop_iter, s);
  return SCIGAP_OK;
}

}  // extern "C"
