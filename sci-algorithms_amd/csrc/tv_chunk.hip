// Chambolle TV of any depth in chunks of up to 6 fused dual updates.
//
// scikit-image's denoise_tv_chambolle (0.18.3 restoration/_denoise.py:341-393)
// runs n_iter_max - 1 dual updates per frame; the legacy solvers use 30
// (PnP_SCI/python/pnp_sci_algo.py:874,895) and scikit-image defaults to 200.
// The tile engine (tv_tile.h) fuses up to 6 updates on a register tile, but
// it starts from p = 0.  Here a launch continues from a stored dual field:
//
//   chunk c covers updates [i0, i0 + nup): it loads img and p_{i0} over the
//   tile (the halo included, so every interior pixel is exact), forms
//   out_{i0} = img + d(p_{i0}), runs nup updates with the energies
//   E_{i0} .. E_{i0+nup-1}, and stores p_{i0+nup} (p is double-buffered
//   across chunks) -- and out on the last chunk;
//   a decision kernel (one workgroup per frame) sums the chunk's per-tile
//   energy slots in a fixed order, stores E_{i0} .. E_{i0+nup-1} in fp64 and
//   applies the exit test of _denoise.py:384-391 for every running frame;
//   a repair kernel (a no-op unless a frame stopped) recomputes, for each
//   frame that stopped at s inside the chunk, out_s from p_{i0} with s - i0
//   updates.  Stopped frames are skipped by later chunks.
//
// A chunk moves img + p in (3 planes) and p out (2 planes) per frame pixel,
// against about 6 planes per update for one launch per update.  The chunk
// depth is a tuning parameter (scigap_tv_chambolle_depth, picked by the
// autotuner); results do not depend on it.

#include "fused.h"

#include <algorithm>

namespace scigap {

namespace {

using namespace fused;

constexpr int kMaxChunk = 6;

struct ChunkArgs {
  const float* img;   // (T,H,W)
  const float* p_in;  // [T][2][N] p_{i0}, or null for i0 = 0
  float* p_out;       // [T][2][N] p_{i0+nup}, or null on the last chunk
  float* out;         // (T,H,W)
  double* energy;     // [T][ne] E_0 .. E_{ne-1} (written by the decision kernels)
  float* eslot;       // [T][D][eslot_stride] per-tile energies of this chunk
  int eslot_stride;   // tiles of the chunk geometry
  int* stop;          // [T] iteration whose image is returned, -1 while running
  int* fix;           // [T] updates to redo for frames that stopped in this chunk, else -1
  int* nfix;          // frames that stopped in this chunk
  int H, W, T, ne, i0, nup, last;
  float weight, tw;
  double eps;
  int tiles_x, ntiles, per_xcd;
};

// Waves per workgroup for a tile geometry of depth D (rows per lane = D).
template <int D> struct ChunkShape {
  static constexpr int BY = D <= 3 ? 16 : (D <= 5 ? 12 : 8);
};

template <int D, int BY>
struct ChunkLds {
  TileLds<BY> eng;
  float2 img[TileGeom<D, BY, 1>::LH * 64];
  float red[kEnergyFrames * BY * EnergySlots<D>::kPerWave];
};

// Engine run from a loaded dual field: out = img + d(p), then nup updates.
// acc[u] receives this lane's part of E_{i0+u} when `accumulate`.
template <int D, int RY, int BY, bool EDGE>
__device__ __forceinline__ void tv_run_from_p(const float2* img, float (&out)[RY][2],
                                              float (&p0)[RY][2], float (&p1)[RY][2], int nup,
                                              TileLds<BY>& lds, const LanePos& lp,
                                              const EdgeMasks& em, float weight, float tw,
                                              float (&acc)[8], bool accumulate) {
#pragma unroll
  for (int i = 0; i < 8; ++i) acc[i] = 0.f;
  // The previous frame ended in tv_div as well: wait until every wave has
  // read its neighbour's p0 row before it is overwritten.
  __syncthreads();
  tv_div<RY, BY, 1>(img, out, p0, p1, lds, lp, accumulate ? &acc[0] : nullptr);
#pragma unroll
  for (int u = 0; u < D; ++u) {
    if (u < nup) {
      tv_grad<RY, BY, 1, EDGE>(out, p0, p1, lds, lp, em, weight, tw,
                               accumulate ? &acc[u] : nullptr);
      tv_div<RY, BY, 1>(img, out, p0, p1, lds, lp,
                        (accumulate && u + 1 < nup) ? &acc[u + 1] : nullptr);
    }
  }
}

// One tile of a chunk (REPAIR: redo the frames listed in `fix`).
template <int D, int BY, bool INSIDE, bool REPAIR>
__device__ __forceinline__ void chunk_tile(const ChunkArgs& a, int tile, ChunkLds<D, BY>& sh) {
  using G = TileGeom<D, BY, 1>;
  constexpr int RY = G::RY;
  const int H = a.H, W = a.W, T = a.T, N = H * W;
  const TilePlace<D, BY, 1> tp(tile, a.tiles_x, H, W);
  const LanePos& lp = tp.lp;
  const int off0 = tp.off0;
  EdgeMasks em;
  if constexpr (!INSIDE) em.init<RY>(lp, H, W);
  float2* const img = sh.img + lp.w * RY * 64 + lp.lane;
  sh.eng.zero_edges();
  if constexpr (!REPAIR)
    for (int i = threadIdx.x; i < kEnergyFrames * BY * EnergySlots<D>::kPerWave; i += blockDim.x)
      sh.red[i] = 0.f;
  for (int t = 0; t < T; ++t) {
    const int nup = REPAIR ? a.fix[t] : (a.stop[t] < 0 ? a.nup : -1);  // uniform
    if (nup >= 0) {
      float v[RY][2], p0[RY][2], p1[RY][2];
      load_plane<RY, INSIDE>(a.img + size_t(t) * N, v, lp, off0, H, W);
      if (a.p_in) {
        const float* pt = a.p_in + size_t(t) * 2 * N;
        load_plane<RY, INSIDE>(pt, p0, lp, off0, H, W);
        load_plane<RY, INSIDE>(pt + N, p1, lp, off0, H, W);
      } else {
#pragma unroll
        for (int j = 0; j < RY; ++j) p0[j][0] = p0[j][1] = p1[j][0] = p1[j][1] = 0.f;
      }
#pragma unroll
      for (int j = 0; j < RY; ++j) img[j * 64] = make_float2(v[j][0], v[j][1]);
      float out[RY][2];
      float acc[8];
      tv_run_from_p<D, RY, BY, !INSIDE>(img, out, p0, p1, nup, sh.eng, lp, em, a.weight, a.tw,
                                        acc, !REPAIR);
      const bool put_p = !REPAIR && a.p_out;
      const bool put_out = REPAIR || a.last;
      float* po = put_p ? a.p_out + size_t(t) * 2 * N : nullptr;
      float* xo = a.out + size_t(t) * N;
#pragma unroll
      for (int j = 0; j < RY; ++j)
#pragma unroll
        for (int k = 0; k < 2; ++k) {
          if (!tp.owns(j, k)) continue;
          const unsigned o = unsigned(off0 + j * W + k);
          if (put_p) {
            st1(po, N, o, p0[j][k]);
            st1(po + N, N, o, p1[j][k]);
          }
          if (put_out) st1(xo, N, o, out[j][k]);
        }
      if constexpr (!REPAIR) {
        float e[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) e[i] = acc[i];
        energy_add<D, BY, 1, 1>(e, tp.lane_owner,
                                sh.red + (t % kEnergyFrames) * BY * EnergySlots<D>::kPerWave, lp);
      }
    }
    if constexpr (!REPAIR) {
      const int tb = t % kEnergyFrames;
      if (tb == kEnergyFrames - 1 || t == T - 1) {
        // Slots of frames skipped in this block hold stale values; those
        // frames have stopped and the decision never reads them.
        const int t0 = t - tb;
        energy_store<D, BY>(sh.red, tb + 1, a.eslot + size_t(t0) * D * a.eslot_stride,
                            a.eslot_stride, tile);
      }
    }
  }
}

template <int D, int BY>
__global__ void __launch_bounds__(64 * BY) k_tv_chunk(ChunkArgs a) {
  __shared__ ChunkLds<D, BY> sh;
  const int tile = tile_of_block(blockIdx.x, a.per_xcd);
  if (tile >= a.ntiles) return;
  if (tile_inside<D, BY, 1>(tile, a.tiles_x, a.H, a.W))
    chunk_tile<D, BY, true, false>(a, tile, sh);
  else
    chunk_tile<D, BY, false, false>(a, tile, sh);
}

template <int D, int BY>
__global__ void __launch_bounds__(64 * BY) k_tv_chunk_fix(ChunkArgs a) {
  __shared__ ChunkLds<D, BY> sh;
  if (*a.nfix == 0) return;  // the common case
  const int tile = tile_of_block(blockIdx.x, a.per_xcd);
  if (tile >= a.ntiles) return;
  chunk_tile<D, BY, false, true>(a, tile, sh);
}

// Energies of this chunk and the exit test (_denoise.py:384-391) over its
// iterations, one workgroup per frame (grid T; D is the chunk geometry's
// depth, slot rows u < nup are this chunk's updates).
template <int D>
__global__ void __launch_bounds__(256) k_chunk_decide(ChunkArgs a) {
  const int t = blockIdx.x;
  if (a.stop[t] >= 0) {  // stopped in an earlier chunk (workgroup-uniform)
    if (threadIdx.x == 0) a.fix[t] = -1;
    return;
  }
  double* e = a.energy + size_t(t) * a.ne;
  const int w = threadIdx.x >> 6;
  for (int u = w; u < a.nup; u += 4) {  // wave-uniform
    const double v = slot_row_sum(a.eslot + (size_t(t) * D + u) * a.eslot_stride,
                                  a.eslot_stride);
    if ((threadIdx.x & 63) == 0) e[a.i0 + u] = v;
  }
  __threadfence_block();
  __syncthreads();
  if (threadIdx.x == 0) {
    int f = -1;
    const double e0 = e[0];
    for (int i = a.i0 > 1 ? a.i0 : 1; i < a.i0 + a.nup; ++i)
      if (fabs(e[i - 1] - e[i]) < a.eps * e0) {
        a.stop[t] = i;
        f = i - a.i0;
        atomicAdd(a.nfix, 1);
        break;
      }
    a.fix[t] = f;
  }
}

__global__ void k_stop_final(const int* stop, int* stop_iter, int T, int ne) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < T) stop_iter[t] = stop[t] < 0 ? ne : stop[t];
}

__global__ void k_fill_int(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <int D>
int chunk_tiles(int H, int W) {
  constexpr int BY = ChunkShape<D>::BY;
  using G = TileGeom<D, BY, 1>;
  return tiles_x_of<D, BY, 1>(W) * ((H + G::OY - 1) / G::OY);
}

// Chunks of up to d updates run on the tile geometry of depth D = d + 1:
// with p loaded, the first divergence already needs p0 from the row above
// and p1 from the column left of the tile, so the top and left halos need
// d + 1 rows/columns (the bottom and right ones d).
template <int D>
int launch_chunk(ChunkArgs a, hipStream_t s) {
  constexpr int BY = ChunkShape<D>::BY;
  using G = TileGeom<D, BY, 1>;
  a.tiles_x = tiles_x_of<D, BY, 1>(a.W);
  a.ntiles = a.tiles_x * ((a.H + G::OY - 1) / G::OY);
  a.per_xcd = (a.ntiles + 7) / 8;
  a.eslot_stride = a.ntiles;
  const unsigned grid = unsigned(a.per_xcd * 8);
  {
    ProfScope ps("tv_chunk", s);
    k_tv_chunk<D, BY><<<grid, 64 * BY, 0, s>>>(a);
  }
  SCIGAP_LAUNCH_CHECK("tv chunk launch");
  {
    ProfScope ps("tv_chunk_decide", s);
    k_chunk_decide<D><<<unsigned(a.T), 256, 0, s>>>(a);
  }
  SCIGAP_LAUNCH_CHECK("tv chunk decision launch");
  {
    ProfScope ps("tv_chunk_fix", s);
    k_tv_chunk_fix<D, BY><<<grid, 64 * BY, 0, s>>>(a);
  }
  SCIGAP_LAUNCH_CHECK("tv chunk repair launch");
  return SCIGAP_OK;
}

int launch_chunk_depth(int d, const ChunkArgs& a, hipStream_t s) {
  switch (d) {
    case 1: return launch_chunk<2>(a, s);
    case 2: return launch_chunk<3>(a, s);
    case 3: return launch_chunk<4>(a, s);
    case 4: return launch_chunk<5>(a, s);
    case 5: return launch_chunk<6>(a, s);
    case 6: return launch_chunk<7>(a, s);
  }
  return fail(SCIGAP_EUNSUPPORTED, "chunk depth %d not instantiated", d);
}

}  // namespace

int max_chunk_depth() { return kMaxChunk; }

// Energy slots of one chunk: the most tiles any chunk geometry has at HxW,
// times (T frames x 7 slot rows, the deepest chunk geometry).
size_t chunk_slot_floats(int64_t H, int64_t W, int64_t T) {
  int most = 0;
  most = std::max(most, chunk_tiles<2>(int(H), int(W)));
  most = std::max(most, chunk_tiles<3>(int(H), int(W)));
  most = std::max(most, chunk_tiles<4>(int(H), int(W)));
  most = std::max(most, chunk_tiles<5>(int(H), int(W)));
  most = std::max(most, chunk_tiles<6>(int(H), int(W)));
  most = std::max(most, chunk_tiles<7>(int(H), int(W)));
  return size_t(T) * (kMaxChunk + 1) * size_t(most);
}

size_t tv_chunked_workspace(int64_t H, int64_t W, int64_t T, int n, int nchunks_max) {
  const size_t N = size_t(H) * W;
  const int ne = n > 1 ? n - 1 : 1;
  return 2 * align256(2 * size_t(T) * N * 4) + align256(size_t(T) * ne * 8) +
         align256(chunk_slot_floats(H, W, T) * 4) + 2 * align256(size_t(T) * 4) +
         align256(size_t(nchunks_max > 0 ? nchunks_max : 1) * 4);
}

// TV of `image` (T,H,W) into `out` with n_iter_max = n >= 2, in chunks of
// `depth` (1..6) updates.
int tv_chunked(const float* image, float* out, int64_t H, int64_t W, int64_t T, float weight,
               float eps, int n, int depth, int* stop_iter, void* ws, hipStream_t s) {
  if (depth < 1 || depth > kMaxChunk) return fail(SCIGAP_EINVAL, "chunk depth %d", depth);
  const size_t N = size_t(H) * W;
  const int ne = n - 1;
  const int nchunks = (ne + depth - 1) / depth;
  char* base = static_cast<char*>(ws);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* q = base + off;
    off += align256(bytes);
    return q;
  };
  float* P[2];
  P[0] = reinterpret_cast<float*>(take(2 * size_t(T) * N * 4));
  P[1] = reinterpret_cast<float*>(take(2 * size_t(T) * N * 4));
  double* energy = reinterpret_cast<double*>(take(size_t(T) * ne * 8));
  float* slots = reinterpret_cast<float*>(take(chunk_slot_floats(H, W, T) * 4));
  int* stop = reinterpret_cast<int*>(take(size_t(T) * 4));
  int* fix = reinterpret_cast<int*>(take(size_t(T) * 4));
  int* nfix = reinterpret_cast<int*>(take(size_t(nchunks) * 4));
  k_fill_int<<<unsigned((T + 255) / 256), 256, 0, s>>>(stop, int(T), -1);
  SCIGAP_LAUNCH_CHECK("stop fill");
  k_fill_int<<<unsigned((nchunks + 255) / 256), 256, 0, s>>>(nfix, nchunks, 0);
  SCIGAP_LAUNCH_CHECK("counter fill");
  ChunkArgs a{};
  a.img = image;
  a.out = out;
  a.energy = energy;
  a.eslot = slots;
  a.stop = stop;
  a.fix = fix;
  a.H = int(H);
  a.W = int(W);
  a.T = int(T);
  a.ne = ne;
  a.weight = weight;
  a.tw = float(0.25 / double(weight));
  a.eps = eps;
  for (int c = 0; c < nchunks; ++c) {
    a.i0 = c * depth;
    a.nup = ne - a.i0 < depth ? ne - a.i0 : depth;
    a.last = c == nchunks - 1;
    a.p_in = c == 0 ? nullptr : P[c & 1];
    a.p_out = a.last ? nullptr : P[(c + 1) & 1];
    a.nfix = nfix + c;
    if (int rc = launch_chunk_depth(depth, a, s)) return rc;
  }
  if (stop_iter) {
    k_stop_final<<<unsigned((T + 255) / 256), 256, 0, s>>>(stop, stop_iter, int(T), ne);
    SCIGAP_LAUNCH_CHECK("stop launch");
  }
  return SCIGAP_OK;
}

}  // namespace scigap
