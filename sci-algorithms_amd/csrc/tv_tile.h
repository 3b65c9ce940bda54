// Register-tiled Chambolle TV engine for gfx950 (device code only).
//
// Arithmetic of one dual update, per frame (scikit-image 0.18.3
// restoration/_denoise.py:349-392, 2-D, tau = 1/4):
//   d   = p0[r-1,c]*[r>0] - p0[r,c] + p1[r,c-1]*[c>0] - p1[r,c]
//   out = image + d                      (out = image before the first update)
//   g0  = out[r+1,c] - out[r,c]  (0 on the last row)
//   g1  = out[r,c+1] - out[r,c]  (0 on the last column)
//   |g| = sqrt(g0^2 + g1^2)
//   p   = (p - tau*g) / (1 + (tau/weight)*|g|)
//   E   = sum d^2 + weight * sum |g|     (per frame; drives the early exit)
//
// The engine holds the dual field scaled by -1/tau = -4, q = -4p:
//   q <- (q + g) / (1 + (tau/weight)*|g|),   out = image - dq/4,
// with dq the same difference stencil applied to q (dq = -4d).  Scaling by
// a power of two commutes with rounding, so every q is exactly -4 times the
// reference's p and out, d^2 = dq^2/16 and the energies are the reference's
// values bit for bit (outside the subnormal range); the dual step then needs
// an add where p needs an FMA, and the first step from q = 0 only a multiply.
//
// Work decomposition.  A workgroup of BY wavefronts owns a load tile of
// 128 columns x BY*RY rows of one frame (RY = MR*D rows per lane, D = number
// of fused updates).  Lane l of wave w holds columns c0+2l, c0+2l+1 of rows
// r0+w*RY .. r0+w*RY+RY-1 in registers (img, out, p0, p1).  Horizontal
// neighbours come from the adjacent lane through DPP wave_shl / wave_shr;
// vertical neighbours come from the lane's own registers except at the wave's
// first/last row, which are exchanged through 2*128 floats of LDS per wave
// (one barrier per exchange).  Every update invalidates one more ring of the
// load tile, so after D updates the result is exact on the tile interior:
// rows [D, BY*RY-D) and columns [HC, 128-HC) with HC = D rounded up to even.
// Outside the image every gradient is masked to zero, which keeps p = 0 there
// and reproduces the reference's boundary handling at the image border; those
// pixels then contribute nothing to the energies.
#pragma once

#include "scigap_internal.h"

namespace scigap {

template <int D, int BY, int MR>
struct TileGeom {
  static constexpr int RY = MR * D;            // rows per lane
  static constexpr int HC = (D + 1) & ~1;      // column halo (even)
  static constexpr int LW = 128;               // load-tile width
  static constexpr int LH = BY * RY;           // load-tile height
  static constexpr int OX = LW - 2 * HC;       // interior width
  static constexpr int OY = LH - 2 * D;        // interior height
  static_assert(MR == 1 || MR == 2, "row multiplier must be 1 or 2");
  static_assert(BY >= 3, "need at least one interior wave");
  static_assert(D >= 1 && D <= 8, "fused depth out of range");
};

// LDS used by the engine: one row of `out` per wave (row w) followed by a
// zero row, and a zero row followed by one row of `p0` per wave (row w+1), so
// that every wave reads its neighbours' rows without a branch: the first
// wave's row above and the last wave's row below are the zero rows.  Those
// two rows hold a halo strip's outermost rows only, whose values the tile
// discards (the first wave's above is also the reference's p = 0 beyond the
// top edge).  zero_edges() writes them; every kernel calls it before the
// engine's first barrier.
template <int BY>
struct TileLds {
  float xo[(BY + 1) * 128];
  float xp[(BY + 1) * 128];
  __device__ void zero_edges() {
    if (threadIdx.x < 64) {
      reinterpret_cast<float2*>(xo)[BY * 64 + threadIdx.x] = make_float2(0.f, 0.f);
      reinterpret_cast<float2*>(xp)[threadIdx.x] = make_float2(0.f, 0.f);
    }
  }
};

// Per-lane view of where its pixels are (for masks and ownership).
struct LanePos {
  int row0;   // global row of register row 0 (wave-uniform)
  int col;    // global column of register column 0
  int w;      // wave index in the workgroup (wave-uniform)
  int lane;   // lane index
};

// Gradient masks for tiles that touch the image border: a forward
// difference is kept where the reference computes one and forced to zero
// elsewhere (last row/column, and every pixel outside the image).  Rows are
// wave-uniform (bit j of rowok0/rowin, held in scalar registers); columns are
// per lane (two floats per kind).
struct EdgeMasks {
  unsigned rowok0 = 0;  // bit j: 0 <= row0+j < H-1   (vertical difference)
  unsigned rowin = 0;   // bit j: 0 <= row0+j < H     (horizontal difference)
  float colin[2];       // 0 <= col+k < W             (vertical difference)
  float colok1[2];      // 0 <= col+k < W-1           (horizontal difference)
  template <int RY>
  __device__ void init(const LanePos& lp, int H, int W) {
    rowok0 = rowin = 0;
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      int r = lp.row0 + j;
      rowok0 |= unsigned(r >= 0 && r < H - 1) << j;
      rowin |= unsigned(r >= 0 && r < H) << j;
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      int c = lp.col + k;
      colin[k] = (c >= 0 && c < W) ? 1.f : 0.f;
      colok1[k] = (c >= 0 && c < W - 1) ? 1.f : 0.f;
    }
  }
};

// Energy partials of one lane are kept in a float[NH*8] array indexed
// [h*8 + u] (update u, strip half h).  With NH = 2, half 0 covers rows
// [0, RY/2) and half 1 rows [RY/2, RY) of the lane's strip (needed when a
// wave holds both halo and interior rows); with NH = 1 one sum per update.

// out = img + d(p), with img read from LDS (row j at img[j*64]); adds sum d^2
// of each strip half h to dacc[h*8] when dacc is non-null.
template <int RY, int BY, int NH>
__device__ __forceinline__ void tv_div(const float2* img, float (&out)[RY][2],
                                       const float (&p0)[RY][2], const float (&p1)[RY][2],
                                       TileLds<BY>& lds, const LanePos& lp, float* dacc) {
  // Publish this wave's last row of p0 for the wave below.
  float2* xp = reinterpret_cast<float2*>(lds.xp);
  xp[(lp.w + 1) * 64 + lp.lane] = make_float2(p0[RY - 1][0], p0[RY - 1][1]);
  // The img slots are this lane's own: read them before the barrier, so that
  // their latency overlaps the wait instead of following it.
  float2 im[RY];
#pragma unroll
  for (int j = 0; j < RY; ++j) im[j] = img[j * 64];
  __syncthreads();
  const float2 above = xp[lp.w * 64 + lp.lane];
  // Rows 1.. first: only row 0 needs the row of the wave above, whose read
  // is still in flight.  The squares are summed in row order afterwards.
  float d0[RY], d1[RY];
#pragma unroll
  for (int jj = 0; jj < RY; ++jj) {
    const int j = jj + 1 < RY ? jj + 1 : 0;
    float up0 = j > 0 ? p0[j - 1][0] : above.x;
    float up1 = j > 0 ? p0[j - 1][1] : above.y;
    float left0 = from_prev_lane(p1[j][1]);
    d0[j] = (up0 - p0[j][0]) + (left0 - p1[j][0]);
    d1[j] = (up1 - p0[j][1]) + (p1[j][0] - p1[j][1]);
    out[j][0] = fmaf(-0.25f, d0[j], im[j].x);
    out[j][1] = fmaf(-0.25f, d1[j], im[j].y);
  }
  float s[NH] = {};
#pragma unroll
  for (int j = 0; j < RY; ++j) {
    const int h = (NH == 2 && j >= RY / 2) ? 1 : 0;
    s[h] = fmaf(d0[j], d0[j], s[h]);
    s[h] = fmaf(d1[j], d1[j], s[h]);
  }
  if (dacc)
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      dacc[h * 8] = fmaf(0.0625f, s[h], dacc[h * 8]);  // sum dq^2 / 16 = sum d^2
      compute_here(dacc[h * 8]);
    }
}

// One dual update q <- (q + g)/(1 + tw |g|) from the current `out` (ZERO:
// q = 0 before it, the first update); adds weight*sum|g| of each strip half h
// to gacc[h*8] when gacc is non-null.
template <int RY, int BY, int NH, bool EDGE, bool ZERO = false>
__device__ __forceinline__ void tv_grad(const float (&out)[RY][2], float (&p0)[RY][2],
                                        float (&p1)[RY][2], TileLds<BY>& lds, const LanePos& lp,
                                        const EdgeMasks& em, float weight, float tw, float* gacc) {
  float2* xo = reinterpret_cast<float2*>(lds.xo);
  xo[lp.w * 64 + lp.lane] = make_float2(out[0][0], out[0][1]);
  __syncthreads();
  const float2 below = xo[(lp.w + 1) * 64 + lp.lane];
  float s[NH] = {};
#pragma unroll
  for (int j = 0; j < RY; ++j) {
    float dn0 = j + 1 < RY ? out[j + 1][0] : below.x;
    float dn1 = j + 1 < RY ? out[j + 1][1] : below.y;
    float right1 = from_next_lane(out[j][0]);
    float g00 = dn0 - out[j][0];
    float g01 = dn1 - out[j][1];
    float g10 = out[j][1] - out[j][0];
    float g11 = right1 - out[j][1];
    if constexpr (EDGE) {
      const bool v_ok = (em.rowok0 >> j) & 1u, h_ok = (em.rowin >> j) & 1u;
      g00 = v_ok ? g00 * em.colin[0] : 0.f;
      g01 = v_ok ? g01 * em.colin[1] : 0.f;
      g10 = h_ok ? g10 * em.colok1[0] : 0.f;
      g11 = h_ok ? g11 * em.colok1[1] : 0.f;
    }
    float n0 = fast_sqrt(fmaf(g00, g00, g10 * g10));
    float n1 = fast_sqrt(fmaf(g01, g01, g11 * g11));
    const int h = (NH == 2 && j >= RY / 2) ? 1 : 0;
    // The first row of a half assigns (0 + x would be an instruction).
    s[h] = j == h * (RY / 2) ? n0 + n1 : s[h] + (n0 + n1);
    float i0 = fast_rcp(fmaf(tw, n0, 1.f));
    float i1 = fast_rcp(fmaf(tw, n1, 1.f));
    p0[j][0] = (ZERO ? g00 : p0[j][0] + g00) * i0;
    p1[j][0] = (ZERO ? g10 : p1[j][0] + g10) * i0;
    p0[j][1] = (ZERO ? g01 : p0[j][1] + g01) * i1;
    p1[j][1] = (ZERO ? g11 : p1[j][1] + g11) * i1;
    // Materialise the new dual values here; otherwise the compiler folds the
    // multiply by 1/den into the next divergence and keeps both factors live.
    asm("" : "+v"(p0[j][0]), "+v"(p1[j][0]), "+v"(p0[j][1]), "+v"(p1[j][1]));
  }
  if (gacc)
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      gacc[h * 8] = fmaf(weight, s[h], gacc[h * 8]);
      compute_here(gacc[h * 8]);
    }
}

// Run `nup` (<= D) updates starting from p = 0 (out = img).  `img` points at
// this lane's float2 of register row 0 in an LDS plane laid out [row][64 lanes]
// (row j at img[j*64]).  On return `out` holds the image after nup updates and
// acc.e[h][u] this lane's partial of E_u over half h of its strip.  When
// `accumulate` is false the energies are not formed (recomputation passes).
// Energy slots are addressed as e[h][u] = base[h*8 + u], so D <= 8.  `hook`
// is called with u after the dual update u (the caller spreads its memory
// requests over the updates with it).
struct NoHook {
  __device__ void operator()(int) const {}
};

template <int D, int RY, int BY, int NH, bool EDGE, class Hook = NoHook>
__device__ __forceinline__ void tv_run_from_zero(const float2* img, float (&out)[RY][2], int nup,
                                                 TileLds<BY>& lds, const LanePos& lp,
                                                 const EdgeMasks& em, float weight, float tw,
                                                 float (&acc)[NH * 8], bool accumulate,
                                                 const Hook& hook = Hook()) {
  float p0[RY][2], p1[RY][2];
#pragma unroll
  for (int j = 0; j < RY; ++j) {
    const float2 im = img[j * 64];
    out[j][0] = im.x;
    out[j][1] = im.y;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      p0[j][k] = 0.f;
      p1[j][k] = 0.f;
    }
  }
#pragma unroll
  for (int i = 0; i < NH * 8; ++i) acc[i] = 0.f;
#pragma unroll
  for (int u = 0; u < D; ++u) {
    if (u < nup) {
      if (u == 0)
        tv_grad<RY, BY, NH, EDGE, true>(out, p0, p1, lds, lp, em, weight, tw,
                                        accumulate ? &acc[u] : nullptr);
      else
        tv_grad<RY, BY, NH, EDGE>(out, p0, p1, lds, lp, em, weight, tw,
                                  accumulate ? &acc[u] : nullptr);
      hook(u);
      const bool dn = accumulate && u + 1 < D;
      tv_div<RY, BY, NH>(img, out, p0, p1, lds, lp, dn ? &acc[u + 1] : nullptr);
    }
  }
}

// One-barrier variant of the engine.  Each update publishes, per wave, the
// new p0 of its last row (the wave below needs it for its first divergence
// row) and the new p0, p1 of its first row, from which the wave above forms
// the first `out` row of this wave itself (the same arithmetic as this wave's
// own divergence, so the values are identical).  One barrier per update
// instead of two; the slots alternate with the update parity so that a slot
// is never rewritten before every reader has passed the next barrier.
template <int BY>
struct TileLdsX1 {
  float2 last0[2][BY * 64];
  float2 first0[2][BY * 64];
  float2 first1[2][BY * 64];
};

// Dual update from `out` with the next wave's first `out` row in `below`
// (no LDS traffic); adds weight*sum|g| per strip half to gacc[h*8].
template <int RY, int NH, bool EDGE, bool ZERO = false>
__device__ __forceinline__ void tv_grad_regs(const float (&out)[RY][2], float2 below,
                                             float (&p0)[RY][2], float (&p1)[RY][2],
                                             const EdgeMasks& em, float weight, float tw,
                                             float* gacc) {
  float s[NH] = {};
#pragma unroll
  for (int j = 0; j < RY; ++j) {
    float dn0 = j + 1 < RY ? out[j + 1][0] : below.x;
    float dn1 = j + 1 < RY ? out[j + 1][1] : below.y;
    float right1 = from_next_lane(out[j][0]);
    float g00 = dn0 - out[j][0];
    float g01 = dn1 - out[j][1];
    float g10 = out[j][1] - out[j][0];
    float g11 = right1 - out[j][1];
    if constexpr (EDGE) {
      const bool v_ok = (em.rowok0 >> j) & 1u, h_ok = (em.rowin >> j) & 1u;
      g00 = v_ok ? g00 * em.colin[0] : 0.f;
      g01 = v_ok ? g01 * em.colin[1] : 0.f;
      g10 = h_ok ? g10 * em.colok1[0] : 0.f;
      g11 = h_ok ? g11 * em.colok1[1] : 0.f;
    }
    float n0 = fast_sqrt(fmaf(g00, g00, g10 * g10));
    float n1 = fast_sqrt(fmaf(g01, g01, g11 * g11));
    const int h = (NH == 2 && j >= RY / 2) ? 1 : 0;
    // The first row of a half assigns (0 + x would be an instruction).
    s[h] = j == h * (RY / 2) ? n0 + n1 : s[h] + (n0 + n1);
    float i0 = fast_rcp(fmaf(tw, n0, 1.f));
    float i1 = fast_rcp(fmaf(tw, n1, 1.f));
    p0[j][0] = (ZERO ? g00 : p0[j][0] + g00) * i0;
    p1[j][0] = (ZERO ? g10 : p1[j][0] + g10) * i0;
    p0[j][1] = (ZERO ? g01 : p0[j][1] + g01) * i1;
    p1[j][1] = (ZERO ? g11 : p1[j][1] + g11) * i1;
    asm("" : "+v"(p0[j][0]), "+v"(p1[j][0]), "+v"(p0[j][1]), "+v"(p1[j][1]));
  }
  if (gacc)
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      gacc[h * 8] = fmaf(weight, s[h], gacc[h * 8]);
      compute_here(gacc[h * 8]);
    }
}

// out = img + d(p) with the p0 row above the strip in `above`.
template <int RY, int NH>
__device__ __forceinline__ void tv_div_regs(const float2* img, float (&out)[RY][2],
                                            const float (&p0)[RY][2], const float (&p1)[RY][2],
                                            float2 above, float* dacc) {
  float s[NH] = {};
#pragma unroll
  for (int j = 0; j < RY; ++j) {
    float up0 = j > 0 ? p0[j - 1][0] : above.x;
    float up1 = j > 0 ? p0[j - 1][1] : above.y;
    float left0 = from_prev_lane(p1[j][1]);
    float d0 = (up0 - p0[j][0]) + (left0 - p1[j][0]);
    float d1 = (up1 - p0[j][1]) + (p1[j][0] - p1[j][1]);
    const float2 im = img[j * 64];
    out[j][0] = fmaf(-0.25f, d0, im.x);
    out[j][1] = fmaf(-0.25f, d1, im.y);
    const int h = (NH == 2 && j >= RY / 2) ? 1 : 0;
    s[h] = fmaf(d0, d0, s[h]);
    s[h] = fmaf(d1, d1, s[h]);
  }
  if (dacc)
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      dacc[h * 8] = fmaf(0.0625f, s[h], dacc[h * 8]);  // sum dq^2 / 16 = sum d^2
      compute_here(dacc[h * 8]);
    }
}

// tv_run_from_zero with one barrier per update (plus one at entry, after
// which the img rows of the wave below are read).  Same arguments and
// results, bit for bit; hook(u) runs after the dual step of update u.
template <int D, int RY, int BY, int NH, bool EDGE, class Hook = NoHook>
__device__ __forceinline__ void tv_run_x1(const float2* img, float (&out)[RY][2], int nup,
                                          TileLdsX1<BY>& lds, const LanePos& lp,
                                          const EdgeMasks& em, float weight, float tw,
                                          float (&acc)[NH * 8], bool accumulate,
                                          const Hook& hook = Hook()) {
  float p0[RY][2], p1[RY][2];
#pragma unroll
  for (int j = 0; j < RY; ++j) {
    const float2 im = img[j * 64];
    out[j][0] = im.x;
    out[j][1] = im.y;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      p0[j][k] = 0.f;
      p1[j][k] = 0.f;
    }
  }
#pragma unroll
  for (int i = 0; i < NH * 8; ++i) acc[i] = 0.f;
  const bool has_below = lp.w + 1 < BY;
  const bool has_above = lp.w > 0;
  __syncthreads();  // img rows of the wave below
  const float2 bimg = has_below ? img[RY * 64] : make_float2(0.f, 0.f);
  float2 below = has_below ? bimg : make_float2(out[RY - 1][0], out[RY - 1][1]);
  const int me = lp.w * 64 + lp.lane;
#pragma unroll
  for (int u = 0; u < D; ++u) {
    if (u < nup) {
      if (u == 0)
        tv_grad_regs<RY, NH, EDGE, true>(out, below, p0, p1, em, weight, tw,
                                         accumulate ? &acc[u] : nullptr);
      else
        tv_grad_regs<RY, NH, EDGE>(out, below, p0, p1, em, weight, tw,
                                   accumulate ? &acc[u] : nullptr);
      hook(u);
      const int par = u & 1;
      lds.last0[par][me] = make_float2(p0[RY - 1][0], p0[RY - 1][1]);
      lds.first0[par][me] = make_float2(p0[0][0], p0[0][1]);
      lds.first1[par][me] = make_float2(p1[0][0], p1[0][1]);
      __syncthreads();
      const float2 above = has_above ? lds.last0[par][me - 64] : make_float2(0.f, 0.f);
      float2 bp0 = make_float2(0.f, 0.f), bp1 = make_float2(0.f, 0.f);
      if (has_below) {
        bp0 = lds.first0[par][me + 64];
        bp1 = lds.first1[par][me + 64];
      }
      const bool dn = accumulate && u + 1 < D;
      tv_div_regs<RY, NH>(img, out, p0, p1, above, dn ? &acc[u + 1] : nullptr);
      // first out row of the wave below, as that wave forms it in tv_div_regs
      const float left0 = from_prev_lane(bp1.y);
      const float d0 = (p0[RY - 1][0] - bp0.x) + (left0 - bp1.x);
      const float d1 = (p0[RY - 1][1] - bp0.y) + (bp1.x - bp1.y);
      below = has_below ? make_float2(fmaf(-0.25f, d0, bimg.x), fmaf(-0.25f, d1, bimg.y))
                        : make_float2(out[RY - 1][0], out[RY - 1][1]);
    }
  }
}

// Frame energies.  Each wave reduces its lanes' partials and writes the
// result to its own LDS slots of the frame ([frame % kEnergyFrames][wave]
// [slot] floats), so frames need no barrier and no atomics; energy_store sums
// the waves of a block of frames in a fixed order and writes one total per
// (frame, update) to the tile's own global slot (one barrier pair per
// workgroup per block, no atomics).  A decision kernel later sums the tiles'
// slots of a frame in a fixed order, so the energies, and with them every
// early-exit decision, do not depend on the order in which workgroups run.
// Halo waves hold no interior rows and never write: their LDS slots keep the
// zeros written at the start of the tile.  A lane contributes half h of its
// strip when those rows are interior rows of the tile and its columns are
// interior columns.
//
// D <= 4 (transposed row sums): within each quad of lanes the D <= 4 partials
// are exchanged so that lane q of the quad ends with the quad's sum of one
// update, e(q) = 2*(q&1) + (q>>1) (three selects-and-adds); two row rotations
// then sum the four quads of each 16-lane row, and lanes 0-3 of every row
// write their update's row sum: 16 slots per wave, [row][update].  D > 4: one
// wave-uniform sum per update (DPP within rows, then lane reads), D slots.
constexpr int kEnergyFrames = 32;  // frames buffered in LDS between flushes

template <int D>
struct EnergySlots {
  static constexpr bool kRows = D <= 4;
  static constexpr int kRowsPerWave = kRows ? 4 : 1;
  static constexpr int kPerWave = kRows ? 16 : D;  // floats per wave and frame
  __host__ __device__ static constexpr int slot(int row, int u) { return kRows ? row * 4 + u : u; }
};

template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xf,
                                                            0xf, true));
}

template <int D, int BY, int MR, int NH>
__device__ __forceinline__ void energy_add(const float (&acc)[NH * 8], bool lane_owner,
                                           float* frame_slots /* LDS, BY * kPerWave floats */,
                                           const LanePos& lp) {
  using ES = EnergySlots<D>;
  bool lo_in, hi_in;
  if (MR == 2) {
    lo_in = lp.w > 0;          // rows [0, D) of wave 0 are the top halo
    hi_in = lp.w < BY - 1;     // rows [D, 2D) of the last wave are the bottom halo
  } else {
    lo_in = hi_in = lp.w > 0 && lp.w < BY - 1;
  }
  if (!lo_in && !hi_in) return;  // wave-uniform: halo waves hold no interior rows
  constexpr int P = ES::kRows ? 4 : D;
  float s[P];
#pragma unroll
  for (int u = 0; u < P; ++u) {
    float v = 0.f;
    if (u < D && lane_owner) {
      v = lo_in ? acc[u] : 0.f;
      if (NH == 2) v += hi_in ? acc[8 + u] : 0.f;
    }
    s[u] = v;
  }
  if constexpr (ES::kRows) {
    const int l = lp.lane;
    const bool b0 = l & 1, b1 = (l >> 1) & 1;
    // Quad lanes 2i and 2i+1 swap halves: even lanes keep updates {0, 1},
    // odd lanes {2, 3} (quad_perm [1,0,3,2]) ...
    float a0 = b0 ? s[2] : s[0], c0 = b0 ? s[0] : s[2];
    float a1 = b0 ? s[3] : s[1], c1 = b0 ? s[1] : s[3];
    a0 += dpp_move<0xB1>(c0);
    a1 += dpp_move<0xB1>(c1);
    // ... then lanes i and i^2 (quad_perm [2,3,0,1]): lane q keeps update
    // 2*(q&1) + (q>>1), summed over the quad.
    float k = b1 ? a1 : a0;
    const float c = b1 ? a0 : a1;
    k += dpp_move<0x4E>(c);
    // The four quads of a row hold the same updates at the same quad
    // positions: row_ror 4, then row_ror 8.
    k += dpp_move<0x124>(k);
    k += dpp_move<0x128>(k);
    const int q = l & 15;
    const int e = 2 * (q & 1) + (q >> 1);
    if (q < 4 && e < D) frame_slots[lp.w * ES::kPerWave + ES::slot(l >> 4, e)] = k;
  } else {
    float r[D];
#pragma unroll
    for (int u = 0; u < D; ++u) r[u] = wave_sum_uniform(s[u]);
    if (lp.lane == 0)
#pragma unroll
      for (int u = 0; u < D; ++u) frame_slots[lp.w * ES::kPerWave + u] = r[u];
  }
}

// Energy of update u of frame t in an LDS block of frames: the sum of every
// wave's slots, in a fixed order.
template <int D, int BY>
__device__ __forceinline__ float energy_block_sum(const float* block, int t, int u) {
  using ES = EnergySlots<D>;
  float s = 0.f;
  for (int w = 0; w < BY; ++w) {
    const float* ws = block + (t * BY + w) * ES::kPerWave;
#pragma unroll
    for (int r = 0; r < ES::kRowsPerWave; ++r) s += ws[ES::slot(r, u)];
  }
  return s;
}

// Per-tile energy slots in global memory: slot (t, u, tile) at
// slots[(t * D + u) * stride + tile], fp32 (a tile's sum is formed in fp32 as
// before; the sum over tiles is taken in fp64 by the decision kernel).
// Store frames [t0, t0+nt) of the LDS block for `tile` (first frame t0; D
// updates each).  All threads of the workgroup call it.
template <int D, int BY>
__device__ __forceinline__ void energy_store(const float* block, int nt, float* slots_t0,
                                             int stride, int tile) {
  __syncthreads();
  for (int i = threadIdx.x; i < nt * D; i += blockDim.x) {
    const int t = i / D, u = i - t * D;
    slots_t0[size_t(i) * stride + tile] = energy_block_sum<D, BY>(block, t, u);
  }
  __syncthreads();
}

// Sum over the tiles of one (frame, update) slot row in a fixed order: lane
// i of the calling wave adds tiles i, i+64, ... in fp64, then a butterfly
// (fixed lane pairing) gives every lane the total.  Wave-uniform result.
__device__ __forceinline__ double slot_row_sum(const float* row, int ntiles) {
  double s = 0.0;
  for (int j = int(threadIdx.x & 63); j < ntiles; j += 64) s += double(row[j]);
  return wave_sum(s);
}

// Early-exit iteration of one frame from its energies E_0..E_{D-1}
// (_denoise.py:384-391): the first i in [1, D) with |E_{i-1} - E_i| < eps*E_0,
// else D (no early exit; the caller returns the image after D updates).
template <int D>
__device__ __forceinline__ int exit_iteration(const double* e, double eps) {
  const double e0 = e[0];
  double prev = e0;
#pragma unroll
  for (int i = 1; i < D; ++i) {
    double ei = e[i];
    if (fabs(prev - ei) < eps * e0) return i;
    prev = ei;
  }
  return D;
}

}  // namespace scigap
