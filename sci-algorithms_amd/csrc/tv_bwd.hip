// Fused reverse pass of the fixed-depth Chambolle TV used by the unrolled
// GAP-TV (one workgroup per tile and frame).
//
// Forward (D updates from p_0 = 0, c = tau/w, tau = 1/4; G the masked
// forward difference, G^T its transpose, scikit-image 0.18.3
// restoration/_denoise.py:349-392 without the energy test):
//   out_u = v + G^T p_u,  g_u = G out_u,  n_u = |g_u|,  s_u = 1/(1 + c n_u),
//   a_u = p_u - tau g_u,  p_{u+1} = a_u s_u,  result out_D.
// Reverse, from xbar = dL/dout_D:
//   vbar = xbar,  pbar_D = G xbar,  and for u = D-1 .. 0:
//     abar = pbar_{u+1} s_u,  sbar = <pbar_{u+1}, a_u>,
//     nbar = -c s_u^2 sbar,   wbar += -n_u s_u^2 sbar * (-tau / w^2),
//     gbar = -tau abar + (nbar / n_u) g_u   (0 where n_u = 0),
//     obar = G^T gbar,  vbar += obar,  pbar_u = abar + G obar.
//
// Nothing but x, Phi, r, xbar (and ybb) and vbar touches memory: the dual field p_u that step u
// needs is recomputed in registers (the same arithmetic as the forward
// engine), from the forward state after D/2 updates, kept in registers, for
// u >= D/2 and from p_0 = 0 otherwise.  Layout: a workgroup of 16 waves holds a
// 128 x 64 load tile of one frame, lane l of wave w columns c0+2l, c0+2l+1 of
// rows r0+4w .. r0+4w+3; the stencil state of a step is in registers, v and
// the running adjoints vbar, pbar in LDS planes (each lane reads back only
// its own slots; one 16-wave workgroup per CU leaves the LDS to it).  Rows
// at wave boundaries are exchanged through LDS (one barrier per exchange,
// two alternating buffers).  Each stencil makes one more ring of the load
// tile stale, 2D rings at most on every side, so the workgroup owns the
// 128-4D x 64-4D interior, and up to the image edge on a side where the tile
// meets it (nothing is stale there: the masks are exact).

#include "fused.h"

namespace scigap {

namespace {

using namespace fused;

constexpr int kBwdRY = 4;   // rows per lane
constexpr int kBwdBY = 16;  // waves per workgroup
constexpr float kTauB = 0.25f;

template <int D>
struct BwdGeom {
  static constexpr int HR = 2 * D;  // halo rows on each side
  static constexpr int HC = 2 * D;  // halo columns on each side (even)
  static constexpr int LW = 128;
  static constexpr int LH = kBwdBY * kBwdRY;
  static constexpr int OX = LW - 2 * HC;
  static constexpr int OY = LH - 2 * HR;
  static_assert(OY > 0 && OX > 0, "tile interior is empty");
};

struct BwdArgs {
  const float* x;     // (T,H,W) x_k of the outer iteration
  const float* phi;   // (T,H,W) mask
  const float* r;     // (H,W) residual r_k the forward used
  float lam;          // lambda_k
  const float* xbar;  // (T,H,W) adjoint of the TV output, or with ybb: vbar of the next
                      // outer iteration, xbar = xbar + ybb * Phi formed on load
  const float* ybb;   // (H,W) or null (see xbar)
  float* vbar;        // (T,H,W) adjoint of the TV input
  double* wbar;       // adjoint of the TV weight (accumulated)
  int H, W, tiles_x, tiles_y;
  float tw;           // tau / w
  double wscale;      // dc/dw = -tau / w^2
};

struct BwdLds {
  float2 vimg[kBwdBY * kBwdRY * 64];  // v, [tile row][lane]
  float2 vb[kBwdBY * kBwdRY * 64];    // vbar
  float2 pb0[kBwdBY * kBwdRY * 64];   // pbar (vertical component)
  float2 pb1[kBwdBY * kBwdRY * 64];   // pbar (horizontal component)
  float2 x[2][(kBwdBY + 2) * 64];     // row exchanges: zero row, wave rows, zero row
  float red[kBwdBY];
};

// Row exchange between neighbouring waves: publish this wave's row `mine`
// and return the one wave w+dir published.  Wave w writes row w+1 of a
// buffer whose first and last rows stay zero, so the first wave's row above
// and the last wave's row below are zeros, read without a branch: above the
// tile's first row that is the reference's p = 0 beyond the top edge (a tile
// starts either there or 2D rows of halo inside the image), and the last
// wave's last row is halo or masked (a tile that ends at the image's last row
// is a border tile, whose gradient mask zeroes that difference), so the row
// below never reaches an owned pixel.  Consecutive exchanges alternate
// buffers, and every exchange has a barrier between its write and its reads,
// so a buffer is rewritten only after all reads of its previous contents.
struct Xchg {
  BwdLds& L;
  int par;
  int w, lane;
  __device__ float2 operator()(float2 mine, int dir) {
    float2* b = L.x[par];
    par ^= 1;
    b[(w + 1) * 64 + lane] = mine;
    __syncthreads();
    return b[(w + 1 + dir) * 64 + lane];
  }
};

// Zero rows of the exchange buffers (before the first exchange's barrier).
__device__ __forceinline__ void zero_exchange_edges(BwdLds& L) {
  if (threadIdx.x < 64) {
#pragma unroll
    for (int par = 0; par < 2; ++par) {
      L.x[par][threadIdx.x] = make_float2(0.f, 0.f);
      L.x[par][(kBwdBY + 1) * 64 + threadIdx.x] = make_float2(0.f, 0.f);
    }
  }
}

using Rows = float[kBwdRY][2];

// Masked forward difference of q: g0 vertical, g1 horizontal (the
// arithmetic and masks of tv_grad).
template <bool EDGE>
__device__ __forceinline__ void gdiff(const Rows& q, Xchg& X, const EdgeMasks& em, Rows& g0,
                                      Rows& g1) {
  constexpr int RY = kBwdRY;
  const float2 below = X(make_float2(q[0][0], q[0][1]), +1);
#pragma unroll
  for (int j = 0; j < RY; ++j) {
    const float dn0 = j + 1 < RY ? q[j + 1][0] : below.x;
    const float dn1 = j + 1 < RY ? q[j + 1][1] : below.y;
    const float right1 = from_next_lane(q[j][0]);
    float a = dn0 - q[j][0], b = dn1 - q[j][1];
    float c = q[j][1] - q[j][0], d = right1 - q[j][1];
    if constexpr (EDGE) {
      const bool v_ok = (em.rowok0 >> j) & 1u, h_ok = (em.rowin >> j) & 1u;
      a = v_ok ? a * em.colin[0] : 0.f;
      b = v_ok ? b * em.colin[1] : 0.f;
      c = h_ok ? c * em.colok1[0] : 0.f;
      d = h_ok ? d * em.colok1[1] : 0.f;
    }
    g0[j][0] = a;
    g0[j][1] = b;
    g1[j][0] = c;
    g1[j][1] = d;
  }
}

// out = base + G^T (a0, a1) (the arithmetic of tv_div); base from LDS rows
// (img) or registers.
__device__ __forceinline__ void gtdiv(const Rows& a0, const Rows& a1, Xchg& X, Rows& out) {
  constexpr int RY = kBwdRY;
  const float2 above = X(make_float2(a0[RY - 1][0], a0[RY - 1][1]), -1);
#pragma unroll
  for (int j = 0; j < RY; ++j) {
    const float up0 = j > 0 ? a0[j - 1][0] : above.x;
    const float up1 = j > 0 ? a0[j - 1][1] : above.y;
    const float left0 = from_prev_lane(a1[j][1]);
    out[j][0] = (up0 - a0[j][0]) + (left0 - a1[j][0]);
    out[j][1] = (up1 - a0[j][1]) + (a1[j][0] - a1[j][1]);
  }
}

// One forward update on the register tile: p <- (p - tau g)/(1 + c|g|),
// g = G out; then out = v + G^T p (v from this lane's LDS slots).
template <bool EDGE>
__device__ __forceinline__ void fwd_update(Rows& out, Rows& p0, Rows& p1, Xchg& X,
                                           const EdgeMasks& em, const float2* vimg, float tw) {
  constexpr int RY = kBwdRY;
  float g0[RY][2], g1[RY][2];
  gdiff<EDGE>(out, X, em, g0, g1);
#pragma unroll
  for (int j = 0; j < RY; ++j)
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const float n = fast_sqrt(fmaf(g0[j][k], g0[j][k], g1[j][k] * g1[j][k]));
      const float s = fast_rcp(fmaf(tw, n, 1.f));
      p0[j][k] = fmaf(-kTauB, g0[j][k], p0[j][k]) * s;
      p1[j][k] = fmaf(-kTauB, g1[j][k], p1[j][k]) * s;
    }
  gtdiv(p0, p1, X, out);
#pragma unroll
  for (int j = 0; j < RY; ++j) {
    const float2 im = vimg[j * 64];
    out[j][0] = im.x + out[j][0];
    out[j][1] = im.y + out[j][1];
  }
}

template <int D, bool INSIDE>
__device__ __forceinline__ void bwd_tile(const BwdArgs& a, int tile, int t, BwdLds& L) {
  using G = BwdGeom<D>;
  constexpr int RY = kBwdRY;
  const int H = a.H, W = a.W, N = H * W;
  const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
  // Edge-aligned placement in both directions (col_place): a tile at the
  // image edge needs no halo on that side and owns up to it.
  const ColPlace cp = col_place(tx, a.tiles_x, W, G::LW, G::HC, G::OX);
  const ColPlace rp = col_place(ty, a.tiles_y, H, G::LH, G::HR, G::OY);
  const int c0 = cp.c0, r0 = rp.c0;
  // Tiles that end at the image's last column or row need its gradient mask.
  const bool inside = c0 >= 0 && r0 >= 0 && c0 + G::LW <= W && r0 + G::LH <= H &&
                      (W % 2 == 0) && !cp.edge_right && !rp.edge_right;
  if (inside != INSIDE) return;
  LanePos lp;
  lp.w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  lp.lane = threadIdx.x & 63;
  lp.row0 = r0 + lp.w * RY;
  lp.col = c0 + 2 * lp.lane;
  const int off0 = lp.row0 * W + lp.col;
  // Owned: the placement's own range, inside the image.
  const bool own_cols = lp.col >= cp.own_lo && lp.col < cp.own_hi;
  const bool own_c[2] = {own_cols && lp.col < W, own_cols && lp.col + 1 < W};
  unsigned own_r = 0;
#pragma unroll
  for (int j = 0; j < RY; ++j) {
    const int row = lp.row0 + j;
    own_r |= unsigned(row >= rp.own_lo && row < rp.own_hi && row < H) << j;
  }
  EdgeMasks em;
  if constexpr (!INSIDE) em.init<RY>(lp, H, W);

  // TV input v = x_k + (lambda_k r_k) Phi, formed as the forward pass forms it.
  // The adjoint of the TV output, xbar = vbar_next + ybb Phi when ybb is
  // given (the frame part of the next projection's reverse), as one FMA.
  float vr[RY][2], xb[RY][2];
  {
    float xv[RY][2], ph[RY][2];
    load_plane<RY, INSIDE>(a.r, vr, lp, off0, H, W);
    load_plane<RY, INSIDE>(a.x + size_t(t) * N, xv, lp, off0, H, W);
    load_plane<RY, INSIDE>(a.phi + size_t(t) * N, ph, lp, off0, H, W);
    load_plane<RY, INSIDE>(a.xbar + size_t(t) * N, xb, lp, off0, H, W);
    if (a.ybb) {
      float yb[RY][2];
      load_plane<RY, INSIDE>(a.ybb, yb, lp, off0, H, W);
#pragma unroll
      for (int j = 0; j < RY; ++j)
#pragma unroll
        for (int k = 0; k < 2; ++k) xb[j][k] = fmaf(yb[j][k], ph[j][k], xb[j][k]);
    }
#pragma unroll
    for (int j = 0; j < RY; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) vr[j][k] = fmaf(vr[j][k] * a.lam, ph[j][k], xv[j][k]);
  }
  // This lane's own slots of the LDS planes (row j at [j * 64]).
  const int slot = lp.w * RY * 64 + lp.lane;
  float2* const vimg = L.vimg + slot;
  float2* const vbs = L.vb + slot;
  float2* const pbs0 = L.pb0 + slot;
  float2* const pbs1 = L.pb1 + slot;
  Xchg X{L, 0, lp.w, lp.lane};
  {
    float g0[RY][2], g1[RY][2];
    gdiff<!INSIDE>(xb, X, em, g0, g1);  // pbar_D = G xbar; vbar = xbar
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      vimg[j * 64] = make_float2(vr[j][0], vr[j][1]);
      vbs[j * 64] = make_float2(xb[j][0], xb[j][1]);
      pbs0[j * 64] = make_float2(g0[j][0], g0[j][1]);
      pbs1[j * 64] = make_float2(g1[j][0], g1[j][1]);
    }
  }
  float cb = 0.f;  // sum of -n s^2 sbar over owned pixels
  const float tw = a.tw;

  // Forward state after M = D/2 updates (p_M, out_M), kept in registers:
  // steps u >= M resume from it instead of from p_0 = 0.
  constexpr int M = D / 2;
  float pm0[RY][2], pm1[RY][2], outm[RY][2];
#pragma unroll
  for (int j = 0; j < RY; ++j) {
    const float2 im = vimg[j * 64];
    outm[j][0] = im.x;
    outm[j][1] = im.y;
    pm0[j][0] = pm0[j][1] = pm1[j][0] = pm1[j][1] = 0.f;
  }
#pragma unroll 1
  for (int i = 0; i < M; ++i) fwd_update<!INSIDE>(outm, pm0, pm1, X, em, vimg, tw);

  // The step and recompute loops stay rolled: unrolled, the compiler
  // interleaves steps and runs out of registers.
#pragma unroll 1
  for (int u = D - 1; u >= 0; --u) {
    // Recompute p_u and out_u: u - M updates from the checkpoint, or u
    // updates from p = 0, out = v.
    const bool resume = u >= M;
    float p0[RY][2], p1[RY][2], out[RY][2];
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      const float2 im = vimg[j * 64];
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        out[j][k] = resume ? outm[j][k] : (k ? im.y : im.x);
        p0[j][k] = resume ? pm0[j][k] : 0.f;
        p1[j][k] = resume ? pm1[j][k] : 0.f;
      }
    }
#pragma unroll 1
    for (int i = resume ? M : 0; i < u; ++i) fwd_update<!INSIDE>(out, p0, p1, X, em, vimg, tw);
    // Step u of the reverse pass.
    float gb0[RY][2], gb1[RY][2];
    {
      float g0[RY][2], g1[RY][2];
      gdiff<!INSIDE>(out, X, em, g0, g1);  // g_u
#pragma unroll
      for (int j = 0; j < RY; ++j) {
        const float2 q0 = pbs0[j * 64], q1 = pbs1[j * 64];
        float ab0[2], ab1[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          const float pb0 = k ? q0.y : q0.x, pb1 = k ? q1.y : q1.x;
          const float n = fast_sqrt(fmaf(g0[j][k], g0[j][k], g1[j][k] * g1[j][k]));
          const float s = fast_rcp(fmaf(tw, n, 1.f));
          const float a0 = fmaf(-kTauB, g0[j][k], p0[j][k]);
          const float a1 = fmaf(-kTauB, g1[j][k], p1[j][k]);
          ab0[k] = pb0 * s;
          ab1[k] = pb1 * s;
          const float sbar = fmaf(pb0, a0, pb1 * a1);
          const float q = s * s * sbar;
          const bool own = own_c[k] && ((own_r >> j) & 1u);
          cb -= own ? n * q : 0.f;
          // nbar / n with nbar = -c q (hardware reciprocal, about 1 ulp)
          const float kk = n > 0.f ? -tw * q * fast_rcp(n) : 0.f;
          gb0[j][k] = fmaf(kk, g0[j][k], -kTauB * ab0[k]);
          gb1[j][k] = fmaf(kk, g1[j][k], -kTauB * ab1[k]);
        }
        pbs0[j * 64] = make_float2(ab0[0], ab0[1]);  // abar, completed to pbar_u below
        pbs1[j * 64] = make_float2(ab1[0], ab1[1]);
      }
    }
    float ob[RY][2];
    gtdiv(gb0, gb1, X, ob);  // obar = G^T gbar
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      const float2 b = vbs[j * 64];
      vbs[j * 64] = make_float2(b.x + ob[j][0], b.y + ob[j][1]);
    }
    if (u > 0) {
      float d0[RY][2], d1[RY][2];
      gdiff<!INSIDE>(ob, X, em, d0, d1);  // pbar_u = abar + G obar
#pragma unroll
      for (int j = 0; j < RY; ++j) {
        const float2 q0 = pbs0[j * 64], q1 = pbs1[j * 64];
        pbs0[j * 64] = make_float2(q0.x + d0[j][0], q0.y + d0[j][1]);
        pbs1[j * 64] = make_float2(q1.x + d1[j][0], q1.y + d1[j][1]);
      }
    }
  }

  float* const dst = a.vbar + size_t(t) * N;
#pragma unroll
  for (int j = 0; j < RY; ++j) {
    if (!((own_r >> j) & 1u)) continue;
    const float2 b = vbs[j * 64];
    if (own_c[0]) st1(dst, N, unsigned(off0 + j * W), b.x);
    if (own_c[1]) st1(dst, N, unsigned(off0 + j * W + 1), b.y);
  }
  // Weight adjoint: one double atomic per workgroup.
  const float ws = wave_sum(cb);
  if (lp.lane == 0) L.red[lp.w] = ws;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int w = 0; w < kBwdBY; ++w) tot += L.red[w];
    atomicAdd(a.wbar, tot * a.wscale);
  }
}

template <int D>
__global__ void __launch_bounds__(64 * kBwdBY) k_tv_bwd(BwdArgs a) {
  __shared__ BwdLds L;
  const int tile = blockIdx.x, t = blockIdx.y;
  zero_exchange_edges(L);
  bwd_tile<D, true>(a, tile, t, L);
  bwd_tile<D, false>(a, tile, t, L);
}

template <int D>
int launch_tv_bwd(BwdArgs a, int T, hipStream_t s) {
  using G = BwdGeom<D>;
  a.tiles_x = tiles_across(a.W, G::LW, G::HC, G::OX);
  a.tiles_y = tiles_across(a.H, G::LH, G::HR, G::OY);
  k_tv_bwd<D><<<dim3(unsigned(a.tiles_x * a.tiles_y), unsigned(T)), 64 * kBwdBY, 0, s>>>(a);
  SCIGAP_LAUNCH_CHECK("fused TV backward launch");
  return SCIGAP_OK;
}

}  // namespace

// vbar = (d out_D / d v)^T xbar for every frame, and wbar += xbar . d out_D / dw,
// for D fixed Chambolle updates with weight w on v = x + (lam r) Phi (D in
// [1, 6]; H*W < 2^30 - 8).  With ybb non-null the adjoint is xbar + ybb*Phi.
int tv_backward_tiles(int D, const float* x, const float* phi, const float* r, float lam,
                      const float* xbar, const float* ybb, float* vbar, int64_t H, int64_t W,
                      int64_t T, float w, double* wbar, hipStream_t s) {
  BwdArgs a{};
  a.x = x;
  a.phi = phi;
  a.r = r;
  a.lam = lam;
  a.xbar = xbar;
  a.ybb = ybb;
  a.vbar = vbar;
  a.wbar = wbar;
  a.H = int(H);
  a.W = int(W);
  a.tw = float(0.25 / double(w));
  a.wscale = -double(kTauB) / (double(w) * double(w));
  switch (D) {
    case 1: return launch_tv_bwd<1>(a, int(T), s);
    case 2: return launch_tv_bwd<2>(a, int(T), s);
    case 3: return launch_tv_bwd<3>(a, int(T), s);
    case 4: return launch_tv_bwd<4>(a, int(T), s);
    case 5: return launch_tv_bwd<5>(a, int(T), s);
    case 6: return launch_tv_bwd<6>(a, int(T), s);
  }
  return fail(SCIGAP_EUNSUPPORTED, "TV backward depth %d not instantiated", D);
}

}  // namespace scigap
