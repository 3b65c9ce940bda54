// Fused GAP/ADMM-TV device code shared by the per-solve (gap_fused.hip) and
// batched (batch.hip) paths: the launch arguments, the fused tile body, the
// early-exit repair kernel and the pointwise kernels of one outer iteration.
// See gap_fused.hip for the algorithm.
#pragma once

#include <type_traits>

#include "tv_tile.h"

namespace scigap {
namespace fused {
// Internal linkage: each translation unit that includes this header gets its
// own copy of these kernels.
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
  float* eslot;         // per-tile energy slots of this iteration, [T][D][eslot_stride]
                        // (tv_tile.h energy_store); null: no energy test, none stored
  int eslot_stride;     // slot row length: >= tiles of the main launch's geometry
  int eslot_tiles;      // tiles of the main launch (slots the decision sums per row)
  float* yb;            // (H,W) sum_t Phi_t*x_t^{new} on owned pixels (for the repair)
  int* stop;            // [T] exit iteration per frame (written by k_decide);
                        // null: no energy test (eps = 0), no decision/repair
  int* nexit;           // frames of this iteration that stopped early (k_decide adds;
                        // zero before the iteration)
  int* stats;           // optional copy of *nexit for the caller (k_decide adds)
  int H, W, T;
  float lam, gam_next, weight, tw;
  double eps;
  int mode, clip, from_y, write_b, read_b, epilogue;
  int fsplit;           // > 1: each tile's frames are split over fsplit workgroups
  float* yb_part;       // [fsplit][H*W] partial yb of each frame part (fsplit > 1)
  int tiles_x, ntiles, per_xcd;
};

// XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs, so
// give XCD i a contiguous band of tiles and let neighbouring tiles (which
// re-read each other's halos) share that XCD's L2.
__device__ __forceinline__ int tile_of_block(int bid, int per_xcd) {
  return (bid & 7) * per_xcd + (bid >> 3);
}

// Buffer-resource access to one (H,W) plane.  A raw buffer access addresses
// base + 32-bit lane offset + uniform (scalar) offset and returns 0 (or drops
// the store) outside [0, N*4) bytes, so a plane needs one VGPR of address per
// lane and an out-of-range offset can never fault.  Planes are < 4 GiB
// (H*W <= 2^30 - 8, checked by the host).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const float* p, int N) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(p), (short)0, unsigned(N) * 4u,
                                           0x00020000);
}
// Two floats per lane in one buffer_load_dwordx2: the LLVM intrinsic itself
// (clang's 64-bit raw buffer-load builtin lowers to a single dword load on
// this toolchain).  Compiler-managed like the builtins: no inline asm.
typedef float buffer_f2 __attribute__((ext_vector_type(2)));
extern "C" __device__ buffer_f2 raw_buffer_load_f2(__amdgpu_buffer_rsrc_t rsrc, int voffset,
                                                   int soffset, int aux)
    __asm("llvm.amdgcn.raw.ptr.buffer.load.v2f32");

__device__ __forceinline__ float ld1(const float* p, int N, unsigned o) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(plane_rsrc(p, N), o * 4u,
                                                                        0, 0));
}
__device__ __forceinline__ void st1(float* p, int N, unsigned o, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), plane_rsrc(p, N), o * 4u, 0, 0);
}

// Load an RY x 2 register block of one plane (zero outside the image).
// `off` is the plane offset of register (0,0).  INSIDE tiles lie entirely in
// the image (and W is even): one 8-byte buffer load per row (a wave reads 512
// contiguous bytes per instruction), the row step in the scalar offset.
// Border tiles send pixels outside the image to an offset past the end of
// the buffer, which reads as 0.
template <int RY, bool INSIDE>
__device__ __forceinline__ void load_plane(const float* __restrict__ src, float (&v)[RY][2],
                                           const LanePos& lp, int off, int H, int W) {
  const __amdgpu_buffer_rsrc_t rs = plane_rsrc(src, H * W);
  if constexpr (INSIDE) {
    const int vo = off * 4;
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      const buffer_f2 f = raw_buffer_load_f2(rs, vo, j * W * 4, 0);
      v[j][0] = f.x;
      v[j][1] = f.y;
    }
  } else {
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      const int r = lp.row0 + j;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const int c = lp.col + k;
        const bool in = r >= 0 && r < H && c >= 0 && c < W;
        const unsigned o = in ? unsigned(off + j * W + k) * 4u : 0xFFFFFFF0u;
        v[j][k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, o, 0, 0));
      }
    }
  }
}

// Row j of load_plane alone.
template <int RY, bool INSIDE>
__device__ __forceinline__ void load_row(const float* __restrict__ src, float (&v)[RY][2],
                                         const LanePos& lp, int off, int H, int W, int j) {
  const __amdgpu_buffer_rsrc_t rs = plane_rsrc(src, H * W);
  if constexpr (INSIDE) {
    const buffer_f2 f = raw_buffer_load_f2(rs, off * 4, j * W * 4, 0);
    v[j][0] = f.x;
    v[j][1] = f.y;
  } else {
    const int r = lp.row0 + j;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int c = lp.col + k;
      const bool in = r >= 0 && r < H && c >= 0 && c < W;
      const unsigned o = in ? unsigned(off + j * W + k) * 4u : 0xFFFFFFF0u;
      v[j][k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, o, 0, 0));
    }
  }
}

// Dual updates over which the requests for the next frame's Phi and x are
// spread (rows of Phi first, then rows of x, evenly).  Issued together
// before the stencil, the eight waves' requests all meet the frame's first
// barrier; spread over three updates the fused launch at 2048x2048x20 takes
// 337 us instead of 369 us (two updates: 344 us, four: 344 us).
constexpr int kPrefetchSpread = 3;

// Column tiling.  When the image is at least one load tile wide (and W is
// even) the first tile starts at column 0 and owns LW-HC columns (no halo is
// needed at the image edge: the missing neighbours are the zeros of the
// reference's boundary), the last tile ends at column W and owns what is
// left (at most LW-HC columns, with a left halo of at least HC), and the
// tiles between own OX columns each.  2048 columns then take 17 tiles instead
// of 18.  Narrower or odd-width images use OX-column tiles from column -HC.
// Ownership boundaries stay even, so every lane owns both or neither of its
// columns.  Host and device code share these functions.
struct ColPlace {
  int c0;           // first loaded column
  int own_lo;       // owned columns [own_lo, own_hi)
  int own_hi;
  bool edge_right;  // ends at the image's last column (needs its gradient mask)
};

__host__ __device__ inline bool cols_aligned(int W, int LW) { return W >= LW && W % 2 == 0; }

__host__ __device__ inline int tiles_across(int W, int LW, int HC, int OX) {
  if (cols_aligned(W, LW)) {
    const int rest = W - 2 * (LW - HC);
    return 2 + (rest > 0 ? (rest + OX - 1) / OX : 0);
  }
  return (W + OX - 1) / OX;
}

__host__ __device__ inline ColPlace col_place(int tx, int tiles_x, int W, int LW, int HC, int OX) {
  ColPlace c;
  if (cols_aligned(W, LW)) {
    if (tx == 0) {
      c.c0 = 0;
      c.own_lo = 0;
      c.own_hi = LW - HC;
    } else if (tx == tiles_x - 1) {
      c.c0 = W - LW;
      c.own_lo = (LW - HC) + (tx - 1) * OX;
      c.own_hi = W;
    } else {
      c.own_lo = (LW - HC) + (tx - 1) * OX;
      c.c0 = c.own_lo - HC;
      c.own_hi = c.own_lo + OX;
    }
    c.edge_right = tx == tiles_x - 1;
  } else {
    c.c0 = tx * OX - HC;
    c.own_lo = c.c0 + HC;
    c.own_hi = c.c0 + LW - HC;
    c.edge_right = false;
  }
  return c;
}

template <int D, int BY, int MR>
__host__ __device__ inline int tiles_x_of(int W) {
  using G = TileGeom<D, BY, MR>;
  return tiles_across(W, G::LW, G::HC, G::OX);
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
    const ColPlace cp = col_place(tx, tiles_x, W, G::LW, G::HC, G::OX);
    const int c0 = cp.c0, r0 = ty * G::OY - D;
    lp.w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    lp.lane = threadIdx.x & 63;
    lp.row0 = r0 + lp.w * G::RY;
    lp.col = c0 + 2 * lp.lane;
    off0 = lp.row0 * W + lp.col;
    inside = c0 >= 0 && r0 >= 0 && c0 + G::LW <= W && r0 + G::LH <= H && (W % 2 == 0) &&
             !cp.edge_right;
    lane_owner = lp.col >= cp.own_lo && lp.col < cp.own_hi;
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
// Registers hold the stencil state (out, p0, p1), lambda*r, the current Phi,
// the yb accumulator and the prefetched next frame (x, Phi); the TV input
// img lives in LDS as a [tile row][128] float plane (the engine reads it on
// every divergence).
template <int LH>
struct DmaPlanes {  // next frame's x and Phi rows, [tile row][128] floats
  float x[LH * 128];
  float ph[LH * 128];
};
struct NoDmaPlanes {};
template <int LH>
struct YbSlots {
  double v[LH * 64 * 2];
};
struct NoYbSlots {};

template <int D, int BY, int MR, bool X1 = false, bool DMA = false>
struct FusedLds {
  static constexpr int LH = TileGeom<D, BY, MR>::LH;
  std::conditional_t<X1, TileLdsX1<BY>, TileLds<BY>> eng;
  float2 img[LH * 64];
  float red[kEnergyFrames * BY * EnergySlots<D>::kPerWave];  // per-frame, per-wave energy partials
  double red64[BY];
  std::conditional_t<DMA, DmaPlanes<LH>, NoDmaPlanes> dma;
  // fp64 A-sum accumulators of the deeper tiles, [tile row][lane][column]
  std::conditional_t<(D > 4), YbSlots<LH>, NoYbSlots> ybs;
};

// Direct-to-LDS prefetch of RY rows (RY even) of one plane for this wave:
// each buffer_load_dwordx4 ... lds moves two rows (lanes 0-31: row j, lanes
// 32-63: row j+1, four floats per lane) into `rows` ([row][128] floats,
// wave-uniform address).  Interior tiles only (the rows lie in the image and
// start 16-byte aligned when W % 4 == 0).
template <int RY>
__device__ __forceinline__ void dma_rows(const float* plane, int N, float* rows, int off0,
                                         int W, int lane) {
  static_assert(RY % 2 == 0, "two rows per transfer");
  const __amdgpu_buffer_rsrc_t rs = plane_rsrc(plane, N);
  // off0 is the offset of this lane's first column pair; the transfer needs
  // the tile row start (lane 0's column) plus 4 floats per lane.
  const int base = off0 - 2 * lane;
  const unsigned vo = unsigned(base + (lane >> 5) * W + 4 * (lane & 31)) * 4u;
#pragma unroll
  for (int j = 0; j < RY; j += 2)
    __builtin_amdgcn_raw_ptr_buffer_load_lds(
        rs, (__attribute__((address_space(3))) void*)(rows + j * 128), 16, vo, j * W * 4, 0, 0);
}

// Stores of the lane's owned pixels (x and, for ADMM, b) of one frame.
// Callers run it under `if (tp.own_c[0])` (per lane); owned rows are
// wave-uniform.  The row offset is formed here, from a base that is made
// opaque once per frame, so that the compiler does not keep RY per-row
// offsets in registers across the frame loop.  ALL_ROWS: the caller knows
// that the wave owns every one of its rows (no per-row tests).
template <int D, int BY, int MR, bool INSIDE, int RY, bool ALL_ROWS = false>
__device__ __forceinline__ void store_owned(float* __restrict__ p, int N, int W,
                                            const TilePlace<D, BY, MR>& tp, unsigned base,
                                            const float (&v)[RY][2]) {
  const __amdgpu_buffer_rsrc_t rs = plane_rsrc(p, N);
#pragma unroll
  for (int j = 0; j < RY; ++j) {
    if constexpr (!ALL_ROWS) {
      if (!((tp.own_r >> j) & 1u)) continue;
    }
    const unsigned o = base + unsigned(j * W) * 4u;
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v[j][0]), rs, o, 0, 0);
    if (INSIDE || tp.own_c[1])
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v[j][1]), rs, o + 4u, 0, 0);
  }
}

// One tile of one fused outer iteration (the body of k_fused and of the
// batched persistent kernel).  Returns without work when the tile's class
// (interior/border) differs from INSIDE.  X1 selects the one-barrier TV
// engine (tv_run_x1).
template <int D, int BY, int MR, int MODE, bool INSIDE, bool IQA, bool X1 = false,
          bool DMA = false>
__device__ __forceinline__ void fused_tile(const FusedArgs& a, int tile,
                                           FusedLds<D, BY, MR, X1, DMA>& sh, int part = 0) {
  using G = TileGeom<D, BY, MR>;
  constexpr int RY = G::RY;
  constexpr int NH = MR;
  // Interior tiles may prefetch the next frame straight into LDS.
  constexpr bool kDma = DMA && INSIDE && (RY % 2 == 0);
  auto& lds = sh.eng;
  if constexpr (!X1) lds.zero_edges();
  float* const red = sh.red;
  double* const red64 = sh.red64;
  const int H = a.H, W = a.W;
  const int N = H * W;
  // Frames [t_lo, t_lo + T) of this work item (all frames unless split).
  const int nparts = a.fsplit > 1 ? a.fsplit : 1;
  const int t_lo = (a.T * part) / nparts;
  const int T = (a.T * (part + 1)) / nparts - t_lo;
  const TilePlace<D, BY, MR> tp(tile, a.tiles_x, H, W);
  if (tp.inside != INSIDE) return;
  const LanePos& lp = tp.lp;
  const int off0 = tp.off0;
  // LDS slot of register row j of this lane: (w*RY + j)*64 + lane.
  const int sl0 = lp.w * RY * 64 + lp.lane;

  EdgeMasks em;
  if constexpr (!INSIDE) em.init<RY>(lp, H, W);

  const bool from_y = a.from_y;
  // q = lambda*r (or y + lambda*r when X_0 = At(y)) over the load tile.
  float q[RY][2];
  // fp64 A-sum accumulators: registers for D <= 4 (115 VGPRs, two workgroups
  // per CU); for the deeper tiles this lane's LDS slots, because fp64
  // registers would cost a resident workgroup (D = 5: 476 -> 662 us per
  // 2048x2048x20 launch).  There an fp32 partial over each block of
  // kYbBlock frames (about kYbBlock ulps) is added to the slot without a
  // return value, once per block.
  constexpr bool kYbInLds = D > 4;
  constexpr int kYbBlock = 8;  // frames per fp32 partial before an LDS add
  double yb[RY][2];
  float ybp[RY][2];  // fp32 partial over the current block of frames
  double* ybl = nullptr;
  if constexpr (kYbInLds) ybl = sh.ybs.v + size_t(sl0) * 2;  // row j at ybl[j * 128]
  if constexpr (MODE != TV_ONLY) {
    load_plane<RY, INSIDE>(a.r_in, q, lp, off0, H, W);
    if (from_y) {
      float yv[RY][2];
      load_plane<RY, INSIDE>(a.y, yv, lp, off0, H, W);
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
  }
  if constexpr (MODE != TV_ONLY) {
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      if constexpr (kYbInLds) {
        ybl[j * 128] = 0.0;
        ybl[j * 128 + 1] = 0.0;
        ybp[j][0] = ybp[j][1] = 0.f;
      } else {
        yb[j][0] = yb[j][1] = 0.0;
      }
    }
  }
  for (int i = threadIdx.x; i < kEnergyFrames * BY * EnergySlots<D>::kPerWave; i += blockDim.x)
    red[i] = 0.f;
  float err = 0.f;

  // Frame t+1's x and Phi are requested before frame t's TV so that their
  // latency hides behind the stencil arithmetic.
  // nx stays 0 when X_0 = At(y) (no previous x), so that img = q*Phi + nx
  // serves both cases.
  float nx[RY][2] = {}, nph[RY][2];
  const float* phit = a.phi + size_t(t_lo) * N;
  const float* xt = a.x_in + size_t(t_lo) * N;
  float* dma_x = nullptr;
  float* dma_ph = nullptr;
  if constexpr (kDma) {
    dma_x = sh.dma.x + lp.w * RY * 128;
    dma_ph = sh.dma.ph + lp.w * RY * 128;
    if constexpr (MODE != TV_ONLY) dma_rows<RY>(phit, N, dma_ph, off0, W, lp.lane);
    if (!from_y) dma_rows<RY>(xt, N, dma_x, off0, W, lp.lane);
  } else {
    if constexpr (MODE != TV_ONLY) load_plane<RY, INSIDE>(phit, nph, lp, off0, H, W);
    if (!from_y) load_plane<RY, INSIDE>(xt, nx, lp, off0, H, W);
  }

  float* xo = a.x_out + size_t(t_lo) * N;
  const float* trt = a.truth ? a.truth + size_t(t_lo) * N : nullptr;
  float* bt = a.b ? a.b + size_t(t_lo) * N : nullptr;
  const int es_stride = a.eslot_stride;
  const int tile_id = tile;
  float* es = a.eslot ? a.eslot + size_t(t_lo) * D * es_stride : nullptr;
  // Every lane reads back only the LDS slots it wrote itself (the one-barrier
  // engine synchronises before it reads a neighbour's rows).
  float2* const img = sh.img + sl0;
  const bool owns_any = tp.own_r != 0;  // wave-uniform: halo waves own nothing
  for (int t = 0; t < T; ++t) {
    float ph[RY][2];
    if constexpr (kDma) {
      // This wave's rows of frame t arrived by DMA (issued one frame ago).
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
#pragma unroll
      for (int j = 0; j < RY; ++j) {
        if constexpr (MODE != TV_ONLY) {
          const float2 f = reinterpret_cast<const float2*>(dma_ph + j * 128)[lp.lane];
          nph[j][0] = f.x;
          nph[j][1] = f.y;
        }
        if (!from_y) {
          const float2 v = reinterpret_cast<const float2*>(dma_x + j * 128)[lp.lane];
          nx[j][0] = v.x;
          nx[j][1] = v.y;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      if constexpr (MODE == TV_ONLY) {
        img[j * 64] = make_float2(nx[j][0], nx[j][1]);
      } else {
#pragma unroll
        for (int k = 0; k < 2; ++k) ph[j][k] = nph[j][k];
        img[j * 64] = make_float2(fmaf(q[j][0], ph[j][0], nx[j][0]),
                                  fmaf(q[j][1], ph[j][1], nx[j][1]));
      }
    }
    if (t + 1 < T) {
      if constexpr (kDma) {
        // The planes were read above; let those reads finish before the
        // next frame's rows overwrite them.
        __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
        if constexpr (MODE != TV_ONLY) dma_rows<RY>(phit + N, N, dma_ph, off0, W, lp.lane);
        if (!from_y) dma_rows<RY>(xt + N, N, dma_x, off0, W, lp.lane);
      }
    }
    // The next frame's Phi and x are requested after the first and second
    // dual updates (their latency then hides behind the rest of this frame's
    // stencil), so that the requests of the eight waves do not all meet the
    // first barrier of the frame.
    // On the last frame the requests read the current frame again (in bounds,
    // and still in the L2 cache), so that no branch guards them.
    const int nstep = t + 1 < T ? N : 0;
    auto prefetch = [&](int u) {
      if constexpr (!kDma) {
        // rows [lo, hi) of a plane go out after update u: Phi first, then x
        constexpr int S = kPrefetchSpread < D ? kPrefetchSpread : D;
        constexpr int total = 2 * RY;
        const int lo = (u * total) / S, hi = ((u + 1) * total) / S;
        if (u >= S) return;
#pragma unroll
        for (int r = 0; r < total; ++r) {
          if (r < lo || r >= hi) continue;
          if (r < RY) {
            if constexpr (MODE != TV_ONLY)
              load_row<RY, INSIDE>(phit + nstep, nph, lp, off0, H, W, r);
          } else if (!from_y) {
            load_row<RY, INSIDE>(xt + nstep, nx, lp, off0, H, W, r - RY);
          }
        }
      }
    };

    float out[RY][2];
    float acc[NH * 8];
    if constexpr (X1) {
      tv_run_x1<D, RY, BY, NH, !INSIDE>(img, out, D, lds, lp, em, a.weight, a.tw, acc, true,
                                        prefetch);
    } else
      tv_run_from_zero<D, RY, BY, NH, !INSIDE>(img, out, D, lds, lp, em, a.weight, a.tw, acc,
                                               true, prefetch);

    if (owns_any) {
      float bv[RY][2];
#pragma unroll
      for (int j = 0; j < RY; ++j) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          float xn = out[j][k];
          if constexpr (MODE == ADMM) {
            const float2 im = img[j * 64];
            const float v = k ? im.y : im.x;
            if (a.clip) xn = fminf(fmaxf(xn, 0.f), 1.f);
            if constexpr (IQA) {  // x_k = v_k + b_k
              if (tp.owns(j, k)) {
                const unsigned o = unsigned(off0 + j * W + k);
                float bk = a.read_b ? ld1(bt, N, o) : 0.f;
                float e = ld1(trt, N, o) - (v + bk);
                err = fmaf(e, e, err);
              }
            }
            bv[j][k] = xn - v;
            if constexpr (kYbInLds)
              ybp[j][k] = fmaf(ph[j][k], 2.f * xn - v, ybp[j][k]);
            else
              yb[j][k] = acc_prod(ph[j][k], 2.f * xn - v, yb[j][k]);
          } else if constexpr (MODE != TV_ONLY) {
            if constexpr (IQA) {
              if (tp.owns(j, k)) {
                float e = ld1(trt, N, unsigned(off0 + j * W + k)) - xn;
                err = fmaf(e, e, err);
              }
            }
            if constexpr (kYbInLds)
              ybp[j][k] = fmaf(ph[j][k], xn, ybp[j][k]);
            else
              yb[j][k] = acc_prod(ph[j][k], xn, yb[j][k]);
          }
          out[j][k] = xn;
        }
      }
      if (tp.own_c[0]) {
        unsigned base = unsigned(off0) * 4u;
        compute_here(base);
        // In an interior tile of one-strip waves (MR = 1) a wave that owns
        // any row owns all of them: strips 1 .. BY-2 are the tile interior.
        constexpr bool kAllRows = INSIDE && MR == 1;
        store_owned<D, BY, MR, INSIDE, RY, kAllRows>(xo, N, W, tp, base, out);
        if constexpr (MODE == ADMM) {
          if (a.write_b) store_owned<D, BY, MR, INSIDE, RY, kAllRows>(bt, N, W, tp, base, bv);
        }
      }
    }
    if constexpr (kYbInLds && MODE != TV_ONLY) {
      if (owns_any && (t % kYbBlock == kYbBlock - 1 || t == T - 1)) {  // wave-uniform
#pragma unroll
        for (int j = 0; j < RY; ++j)
#pragma unroll
          for (int k = 0; k < 2; ++k) {
            atomicAdd(&ybl[j * 128 + k], double(ybp[j][k]));
            ybp[j][k] = 0.f;
          }
      }
    }
    const int tb = t % kEnergyFrames;
    energy_add<D, BY, MR, NH>(acc, tp.lane_owner, red + tb * BY * EnergySlots<D>::kPerWave, lp);
    if (es && (tb == kEnergyFrames - 1 || t == T - 1))  // workgroup-uniform
      energy_store<D, BY>(red, tb + 1, es - size_t(tb) * D * es_stride, es_stride, tile_id);
    phit += N;
    xt += N;
    xo += N;
    if (es) es += size_t(D) * es_stride;
    if constexpr (IQA) trt += N;
    if constexpr (MODE == ADMM) bt += (bt ? N : 0);
  }

  if constexpr (MODE != TV_ONLY) {
#pragma unroll
    for (int j = 0; j < RY; ++j) {
      if (!((tp.own_r >> j) & 1u)) continue;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (!tp.own_c[k]) continue;
        const unsigned o = unsigned(off0 + j * W + k);
        const float v = kYbInLds ? float(ybl[j * 128 + k]) : float(yb[j][k]);
        if (nparts > 1) {
          // the parts are summed (and the epilogue run) by k_part_epilogue
          st1(a.yb_part + size_t(part) * N, N, o, v);
          continue;
        }
        if (a.yb) st1(a.yb, N, o, v);
        if (a.epilogue) epilogue_pixel<MODE>(a, o, v);
      }
    }
  }
  if constexpr (IQA) commit_err<BY>(err, a.err, red64, lp.w, lp.lane);
}

// Sum of the frame parts' yb and the residual epilogue (fsplit > 1): the
// frame loop of one tile is split over workgroups to even out the last
// round of workgroups on the CUs; their partial yb planes meet here.
template <int MODE>
__device__ __forceinline__ void part_epilogue(const FusedArgs& a, int first, int stride) {
  const int N = a.H * a.W;
  for (int i = first; i < N; i += stride) {
    double sum = a.yb_part[i];
    for (int f = 1; f < a.fsplit; ++f) sum += a.yb_part[size_t(f) * N + i];
    const float yb = float(sum);
    if (a.yb) a.yb[i] = yb;
    if (a.epilogue) epilogue_pixel<MODE>(a, unsigned(i), yb);
  }
}

template <int MODE>
__global__ void __launch_bounds__(256) k_part_epilogue(FusedArgs a) {
  part_epilogue<MODE>(a, int(blockIdx.x * blockDim.x + threadIdx.x), int(gridDim.x * blockDim.x));
}

// Whether a tile's load region lies entirely in the image (mirrors TilePlace).
template <int D, int BY, int MR>
__host__ __device__ inline bool tile_inside(int tile, int tiles_x, int H, int W) {
  using G = TileGeom<D, BY, MR>;
  const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
  const ColPlace cp = col_place(tx, tiles_x, W, G::LW, G::HC, G::OX);
  const int c0 = cp.c0, r0 = ty * G::OY - D;
  return c0 >= 0 && r0 >= 0 && c0 + G::LW <= W && r0 + G::LH <= H && (W % 2 == 0) &&
         !cp.edge_right;
}

// One launch covers interior and border tiles; the two bodies are separate
// instantiations (the interior one has no masks or bounds checks).  PART
// restricts a launch to one class (1: interior, 2: border; 0: both), so that
// the interior kernel is allocated registers for its own body only.  With the
// per-iteration error (IQA) the 8-wave kernel is held to 128 VGPRs, the limit
// for two workgroups per CU (it would take 130-133; with one workgroup per CU
// a 2048x2048x20 solve takes 48 ms instead of 32).
template <int D, int BY, int MR, int MODE, bool IQA, bool X1 = false, int PART = 0,
          bool DMA = false>
__global__ void __launch_bounds__(64 * BY)
__attribute__((amdgpu_waves_per_eu(BY == 8 && D <= 4 && IQA ? 4 : 1))) k_fused(FusedArgs a) {
  __shared__ FusedLds<D, BY, MR, X1, DMA> sh;
  // Work item -> (tile, frame part); the parts of a tile are adjacent, so
  // they run on the same XCD.
  const int item = tile_of_block(blockIdx.x, a.per_xcd);
  if (item >= a.ntiles) return;
  const int nparts = a.fsplit > 1 ? a.fsplit : 1;
  const int tile = item / nparts, part = item - tile * nparts;
  const bool inside = tile_inside<D, BY, MR>(tile, a.tiles_x, a.H, a.W);
  if constexpr (PART != 2) {
    if (inside) {
      fused_tile<D, BY, MR, MODE, true, IQA, X1, DMA>(a, tile, sh, part);
      return;
    }
  }
  if constexpr (PART != 1) {
    if (!inside) fused_tile<D, BY, MR, MODE, false, IQA, X1>(a, tile, sh, part);
  }
}

// Early-exit decision of one frame (workgroup blockIdx.x; NT threads): the
// energies E_0..E_{D-1} are the sums of the frame's per-tile slots, taken in
// a fixed order, then the exit test of _denoise.py:384-391.  Thread i adds
// tiles i, i+NT, ... of every update row with all of its loads issued
// together (one memory latency per frame, not one per tile), in fp64; the
// waves then reduce by a fixed butterfly and thread 0 adds the wave sums in
// order.  Writes stop[t] and adds early stops to *nexit and *stats.
template <int D, int NT = 256>
__device__ __forceinline__ void decide_frame(const FusedArgs& a, int t) {
  constexpr int NW = NT / 64;
  __shared__ double part[NW][D];
  const float* rows = a.eslot + size_t(t) * D * a.eslot_stride;
  double acc[D];
#pragma unroll
  for (int u = 0; u < D; ++u) acc[u] = 0.0;
  constexpr int kUnroll = NT >= 1024 ? 2 : 4;
  for (int j0 = int(threadIdx.x); j0 < a.eslot_tiles; j0 += NT * kUnroll) {
    float v[kUnroll][D];
#pragma unroll
    for (int q = 0; q < kUnroll; ++q) {
      const int j = j0 + q * NT;
#pragma unroll
      for (int u = 0; u < D; ++u) v[q][u] = j < a.eslot_tiles ? rows[size_t(u) * a.eslot_stride + j] : 0.f;
    }
#pragma unroll
    for (int q = 0; q < kUnroll; ++q)
#pragma unroll
      for (int u = 0; u < D; ++u) acc[u] += double(v[q][u]);
  }
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int u = 0; u < D; ++u) {
    const double sum = wave_sum(acc[u]);
    if ((threadIdx.x & 63) == 0) part[w][u] = sum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double e[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
      e[u] = part[0][u];
      for (int i = 1; i < NW; ++i) e[u] += part[i][u];
    }
    const int st = exit_iteration<D>(e, a.eps);
    a.stop[t] = st;
    if (st < D) {
      atomicAdd(a.nexit, 1);
      if (a.stats) atomicAdd(a.stats, 1);
    }
  }
}

// Threads of the single-solve decision kernel (one workgroup per frame): at
// 2048x2048x20 (1462 tiles) 1024 threads issue every slot load of a frame in
// one round, 7.2 us per launch against 8.8 us for 512 and 11.8 us for 256.
constexpr int kDecideThreads = 1024;

template <int D>
__global__ void __launch_bounds__(kDecideThreads) k_decide(FusedArgs a) {
  decide_frame<D, kDecideThreads>(a, blockIdx.x);
}

// The decision and, for a main pass split into frame parts, the parts' yb sum
// and residual epilogue in one launch: workgroups [0, T) decide one frame
// each, the rest run the epilogue over the pixels (the two read different
// results of the main pass, and the repair after them needs both).  One
// launch less per outer iteration where launches dominate (small problems).
template <int D, int MODE>
__global__ void __launch_bounds__(kDecideThreads) k_decide_epilogue(FusedArgs a) {
  if (blockIdx.x < unsigned(a.T)) {  // workgroup-uniform
    decide_frame<D, kDecideThreads>(a, blockIdx.x);
    return;
  }
  const int b = int(blockIdx.x) - a.T, nb = int(gridDim.x) - a.T;
  part_epilogue<MODE>(a, b * kDecideThreads + int(threadIdx.x), nb * kDecideThreads);
}

// Early-exit repair (does work only when some frame stopped early; `stop`
// holds each frame's exit iteration, D for none).
// For every frame that stopped at s < D, the frame is recomputed with s
// updates from the unchanged inputs (x_in and r_in are double-buffered), its
// contribution to yb is corrected by Phi*(x_new - x_old) (ADMM: 2*Phi*...),
// and the residual epilogue is recomputed from the corrected yb.
template <int D, int BY, int MR>
struct FixLds {
  TileLds<BY> eng;
  float2 img[TileGeom<D, BY, MR>::LH * 64];
  double red64[BY];
};

template <int D, int BY, int MR, int MODE, bool INSIDE>
__device__ __forceinline__ void fused_fix_tile(const FusedArgs& a, int tile, FixLds<D, BY, MR>& sh,
                                               const int* __restrict__ stop) {
  using G = TileGeom<D, BY, MR>;
  constexpr int RY = G::RY;
  constexpr int NH = MR;
  constexpr bool kTvOnly = MODE == TV_ONLY;
  TileLds<BY>& lds = sh.eng;
  lds.zero_edges();
  double* const red64 = sh.red64;
  const int H = a.H, W = a.W, T = a.T;
  const int N = H * W;
  const TilePlace<D, BY, MR> tp(tile, a.tiles_x, H, W);
  const LanePos& lp = tp.lp;
  const int off0 = tp.off0;
  EdgeMasks em;
  if constexpr (!INSIDE) em.init<RY>(lp, H, W);
  const bool from_y = a.from_y;

  // q = lambda*r (or y + lambda*r when X_0 = At(y)), as in the main pass.
  float q[RY][2] = {};
  if constexpr (!kTvOnly) {
    load_plane<RY, INSIDE>(a.r_in, q, lp, off0, H, W);
    if (from_y) {
      float yv[RY][2];
      load_plane<RY, INSIDE>(a.y, yv, lp, off0, H, W);
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
  }
  float dyb[RY][2] = {};
  float err = 0.f;

  // The stopped frames are visited in order.  A frame's x and Phi are
  // requested one stopped frame ahead and its old output before its TV, so
  // that their latency hides behind the stencil arithmetic (as in the main
  // pass); nx stays 0 when X_0 = At(y).
  // The exit iterations come from LDS (or memory) into VGPRs: readfirstlane
  // tells the compiler that they, the frame indices and so the plane
  // pointers are wave-uniform (otherwise every buffer load of a frame is
  // wrapped in a loop over the lanes' distinct descriptors).
  auto next_stopped = [&](int t) {
    while (t < T && __builtin_amdgcn_readfirstlane(stop[t]) >= D) ++t;
    return t;
  };
  int t = next_stopped(0);
  float nph[RY][2] = {}, nx[RY][2] = {};
  if (t < T) {
    if constexpr (!kTvOnly) load_plane<RY, INSIDE>(a.phi + size_t(t) * N, nph, lp, off0, H, W);
    if (!from_y) load_plane<RY, INSIDE>(a.x_in + size_t(t) * N, nx, lp, off0, H, W);
  }
  // Every lane reads back only the LDS slots it wrote itself.
  float2* const simg = sh.img + lp.w * RY * 64 + lp.lane;
  while (t < T) {
    const int nup = __builtin_amdgcn_readfirstlane(stop[t]);
    float ph[RY][2];
#pragma unroll
    for (int j = 0; j < RY; ++j) {
#pragma unroll
      for (int k = 0; k < 2; ++k) ph[j][k] = nph[j][k];
      simg[j * 64] = kTvOnly ? make_float2(nx[j][0], nx[j][1])
                             : make_float2(fmaf(q[j][0], ph[j][0], nx[j][0]),
                                           fmaf(q[j][1], ph[j][1], nx[j][1]));
    }
    float* xo = a.x_out + size_t(t) * N;
    float xold[RY][2] = {};
    const int tn = next_stopped(t + 1);
    // This frame's old output and the next stopped frame's Phi and x are
    // requested row by row over the frame's dual updates (as in the main
    // pass), so that the waves' requests do not all meet the first barrier.
    const float* nphi = a.phi + size_t(tn < T ? tn : t) * N;
    const float* nxin = a.x_in + size_t(tn < T ? tn : t) * N;
    const int spread = nup < 3 ? nup : 3;
    auto requests = [&](int u) {
      constexpr int total = 3 * RY;
      const int lo = (u * total) / spread, hi = ((u + 1) * total) / spread;
#pragma unroll
      for (int r = 0; r < total; ++r) {
        if (r < lo || r >= hi) continue;
        if (r < RY) {
          if constexpr (!kTvOnly) load_row<RY, INSIDE>(xo, xold, lp, off0, H, W, r);
        } else if (r < 2 * RY) {
          if constexpr (!kTvOnly)
            if (tn < T) load_row<RY, INSIDE>(nphi, nph, lp, off0, H, W, r - RY);
        } else if (!from_y && tn < T) {
          load_row<RY, INSIDE>(nxin, nx, lp, off0, H, W, r - 2 * RY);
        }
      }
    };
    float out[RY][2];
    float acc[NH * 8];
    tv_run_from_zero<D, RY, BY, NH, !INSIDE>(simg, out, nup, lds, lp, em, a.weight, a.tw, acc,
                                             false, requests);
    const float* trt = a.truth ? a.truth + size_t(t) * N : nullptr;
    float* bt = a.b ? a.b + size_t(t) * N : nullptr;
    float bv[RY][2];
#pragma unroll
    for (int j = 0; j < RY; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        float xn = out[j][k];
        if constexpr (MODE == ADMM) {
          if (a.clip) xn = fminf(fmaxf(xn, 0.f), 1.f);
          const float2 im = simg[j * 64];  // this lane's own slot
          bv[j][k] = xn - (k ? im.y : im.x);
          if (tp.owns(j, k)) dyb[j][k] = fmaf(2.f * ph[j][k], xn - xold[j][k], dyb[j][k]);
        } else if constexpr (!kTvOnly) {
          if (tp.owns(j, k)) {
            if (trt) {  // replace the main pass's error of this pixel
              const unsigned o = unsigned(off0 + j * W + k);
              float e = trt[o] - xn, eo = trt[o] - xold[j][k];
              err = fmaf(e, e, err);
              err = fmaf(-eo, eo, err);
            }
            dyb[j][k] = fmaf(ph[j][k], xn - xold[j][k], dyb[j][k]);
          }
        }
        out[j][k] = xn;
      }
    // Owned pixels only, one 8-byte store per row where both columns are owned.
    if (tp.own_c[0]) {
      unsigned base = unsigned(off0) * 4u;
      compute_here(base);
      store_owned<D, BY, MR, INSIDE>(xo, N, W, tp, base, out);
      if constexpr (MODE == ADMM)
        if (a.write_b) store_owned<D, BY, MR, INSIDE>(bt, N, W, tp, base, bv);
    }
    t = tn;
  }
  if constexpr (!kTvOnly) {
#pragma unroll
    for (int j = 0; j < RY; ++j)
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        if (!tp.owns(j, k)) continue;
        const unsigned o = unsigned(off0 + j * W + k);
        // The corrections are small differences: their fp32 sum is exact enough.
        const float yb = a.yb[o] + dyb[j][k];
        a.yb[o] = yb;
        if (a.epilogue) epilogue_pixel<MODE>(a, o, yb);
      }
  }
  if constexpr (!kTvOnly && MODE != ADMM)
    if (a.truth && a.err) commit_err<BY>(err, a.err, red64, lp.w, lp.lane);
}

// Frames whose exit iterations the repair kernel copies into LDS (the frame
// scans then cost no memory latency); longer solves read them from memory.
constexpr int kStopLds = 1024;

// For D <= 4 at least 4 waves per SIMD: two 8-wave workgroups per CU
// (128 VGPRs); deeper tiles need more registers than that.  Runs after
// k_decide; returns at once when no frame stopped early (the common case).
template <int D, int BY, int MR, int MODE>
__global__ void __launch_bounds__(64 * BY) __attribute__((amdgpu_waves_per_eu(D <= 4 ? 4 : 1)))
k_fused_fix(FusedArgs a) {
  __shared__ FixLds<D, BY, MR> sh;
  __shared__ int s_stop[kStopLds];
  if (*a.nexit == 0) return;  // one scalar load per workgroup
  const int tile = tile_of_block(blockIdx.x, a.per_xcd);
  if (tile >= a.ntiles) return;
  const int* stop = a.stop;
  if (a.T <= kStopLds) {
    for (int t = threadIdx.x; t < a.T; t += blockDim.x) s_stop[t] = a.stop[t];
    __syncthreads();
    stop = s_stop;
  }
  // Interior tiles run without masks or bounds checks (as in the main pass).
  if (tile_inside<D, BY, MR>(tile, a.tiles_x, a.H, a.W))
    fused_fix_tile<D, BY, MR, MODE, true>(a, tile, sh, stop);
  else
    fused_fix_tile<D, BY, MR, MODE, false>(a, tile, sh, stop);
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
    const float yv = y[i];
    double sum = 0.0;
    for (int t = 0; t < T; ++t) {
      float f = phi[t * N + i];
      float xv = x0 ? x0[t * N + i] : yv * f;
      sum = acc_prod(xv, f, sum);
    }
    const float yb = float(sum);
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
    double yb = 0.0;
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
        yb = acc_prod(f, 2.f * xn - v, yb);
      } else {
        if (a.truth) {
          float e = a.truth[o] - xn;
          err = fmaf(e, e, err);
        }
        yb = acc_prod(f, xn, yb);
      }
      a.x_out[o] = xn;
    }
    if (a.epilogue) {
      if (a.mode == GAP_ACC) epilogue_pixel<GAP_ACC>(a, unsigned(i), float(yb));
      else if (a.mode == GAP_PLAIN) epilogue_pixel<GAP_PLAIN>(a, unsigned(i), float(yb));
      else epilogue_pixel<ADMM>(a, unsigned(i), float(yb));
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
    double yb = 0.0;
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
        yb = acc_prod(f, 2.f * xn - v, yb);
      } else {
        if (a.truth) {
          float e = a.truth[o] - xn;
          err = fmaf(e, e, err);
        }
        yb = acc_prod(f, xn, yb);
      }
    }
    if (a.epilogue) {
      if (a.mode == GAP_ACC) epilogue_pixel<GAP_ACC>(a, unsigned(i), float(yb));
      else if (a.mode == GAP_PLAIN) epilogue_pixel<GAP_PLAIN>(a, unsigned(i), float(yb));
      else epilogue_pixel<ADMM>(a, unsigned(i), float(yb));
    }
  }
  if (a.truth && a.err) {
    double s = wave_sum(double(err));
    if ((threadIdx.x & 63) == 0) atomicAdd(a.err, s);
  }
}

// *out += number of frames whose TV returned before its last iteration
// (stop[t] < last); integer sums, so the count does not depend on order.
__global__ void __launch_bounds__(256) k_count_early(const int* __restrict__ stop, int T,
                                                     int last, int* out) {
  int c = 0;
  for (int t = int(threadIdx.x); t < T; t += blockDim.x) c += stop[t] < last ? 1 : 0;
  if (c) atomicAdd(out, c);
}

__global__ void k_fill_i32(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}


// ---------------------------------------------------------------------------
// Host side: per-iteration launch arguments (shared by the single-solve and
// batched drivers)
// ---------------------------------------------------------------------------

// Device buffers of one solve.
struct SolveState {
  float* xbuf = nullptr;     // (T,H,W) state buffer paired with x_out
  float* bbuf = nullptr;     // (T,H,W) ADMM b
  float* r[2] = {nullptr, nullptr};
  float* y1[2] = {nullptr, nullptr};
  float* yb = nullptr;       // (H,W)
  float* yb_part = nullptr;  // [parts][H*W] partial yb of frame-split launches
  int* stop = nullptr;       // [T]
  int* nexit = nullptr;      // [iters]
  float* eslot = nullptr;    // [T][D][eslot_stride] per-tile energies (reused every iteration)
  int eslot_stride = 0;
};

inline int solver_mode(const scigap_params& p) {
  return p.method == SCIGAP_METHOD_ADMM ? ADMM : (p.accelerate ? GAP_ACC : GAP_PLAIN);
}

// Outer iterations that run the fused kernel: ADMM's last iteration is a
// pointwise projection (its TV never reaches the returned x).
inline int fused_iterations(const scigap_params& p) {
  return solver_mode(p) == ADMM ? p.iters - 1 : p.iters;
}

inline double decayed(double v, double f, int k) {
  for (int i = 0; i < k; ++i) v *= f;  // the reference multiplies once per iteration
  return v;
}

// Buffer holding state X_j (GAP: x, ADMM: theta) so that the last state
// lands in x_out; X_0 is the caller's x0 (or At(y), formed on the fly).
inline float* state_buffer(const scigap_params& p, const SolveState& st, float* x_out, int j) {
  return ((fused_iterations(p) - j) % 2 == 0) ? x_out : st.xbuf;
}

inline FusedArgs iteration_args(const scigap_params& p, int k, int64_t H, int64_t W, int64_t T,
                                 const float* y, const float* phi, const float* phi_sum,
                                 const float* x0, float* x_out, const float* truth,
                                 double* sq_err, int* stats, const SolveState& st) {
  const int K = p.iters, mode = solver_mode(p), D = p.tv_iter_max - 1;
  const int last_fused = fused_iterations(p) - 1;
  FusedArgs a{};
  a.from_y = (k == 0 && !x0);
  a.x_in = k == 0 ? x0 : state_buffer(p, st, x_out, k);
  a.x_out = state_buffer(p, st, x_out, k + 1);
  a.phi = phi;
  a.y = y;
  a.phi_sum = phi_sum;
  a.r_in = st.r[k & 1];
  a.r_out = st.r[(k + 1) & 1];
  a.y1_in = st.y1[k & 1];
  a.y1_out = st.y1[(k + 1) & 1];
  a.truth = truth;
  a.err = truth ? sq_err + k : nullptr;
  a.b = st.bbuf;
  a.eslot = st.eslot;
  a.eslot_stride = st.eslot_stride;
  a.yb = st.yb;
  a.yb_part = st.yb_part;
  a.stop = st.stop;
  a.nexit = st.nexit + k;
  a.stats = stats ? stats + k : nullptr;
  a.H = int(H);
  a.W = int(W);
  a.T = int(T);
  a.lam = p.lambda_;
  a.gam_next = float(decayed(p.gamma, p.gamma_decay, k + 1));
  const double wk = decayed(p.tv_weight, p.tv_weight_decay, k);
  a.weight = float(wk);
  a.tw = float(0.25 / wk);
  a.eps = p.tv_eps;
  a.mode = mode;
  a.clip = p.clip;
  a.write_b = mode == ADMM && (truth != nullptr || k == last_fused);
  a.read_b = mode == ADMM && truth != nullptr && k > 0;
  a.epilogue = (k + 1 < K);
  return a;
}

}  // namespace
}  // namespace fused
}  // namespace scigap
