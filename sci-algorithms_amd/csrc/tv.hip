// General-depth Chambolle TV (any n_iter_max): one launch per dual update.
//
// Used when n_iter_max-1 exceeds the depth the fused tile engine handles
// (tv_tile.h), e.g. the 30 iterations of the legacy solvers
// (PnP_SCI/python/pnp_sci_algo.py:874,895) or scikit-image's default 200.
// Launch u (0 <= u < n-1) reads p_u (p_0 = 0), forms out_u = img + d(p_u),
// accumulates E_u per frame, writes p_{u+1} and (u >= 1) out_u; launch n-1
// writes out_{n-1}.  Before any work a frame checks the exit test of
// _denoise.py:384-391 on E_0..E_{u-1}; once it has stopped at s, out_s is
// already in place and later launches leave the frame alone, so the early
// exit is exact without a repair pass.  p is double-buffered across launches
// because each launch reads its neighbours' p_u.

#include "scigap_internal.h"

namespace scigap {

int fill_f64(double* p, int n, double v, hipStream_t s);

namespace {

constexpr int kBX = 64, kBY = 4;

struct StepArgs {
  const float* img;   // (T,H,W)
  const float* p_in;  // [T][2][H*W] p_u, or null for u == 0
  float* p_out;       // [T][2][H*W] p_{u+1}, or null on the last launch
  float* out;         // (T,H,W)
  double* energy;     // [T][n]
  int H, W, T, n, u;
  float weight, tw;
  double eps;
};

// First iteration i in [1, upto) at which the frame stopped, or -1.
__device__ __forceinline__ int stopped_at(const double* e, int upto, double eps) {
  const double e0 = e[0];
  for (int i = 1; i < upto; ++i)
    if (fabs(e[i - 1] - e[i]) < eps * e0) return i;
  return -1;
}

__global__ void __launch_bounds__(kBX* kBY) k_tv_step(StepArgs a) {
  const int t = blockIdx.z;
  const int H = a.H, W = a.W;
  const size_t N = size_t(H) * W;
  if (stopped_at(a.energy + size_t(t) * a.n, a.u, a.eps) >= 0) return;  // uniform

  const int c = blockIdx.x * kBX + (threadIdx.x % kBX);
  const int r = blockIdx.y * kBY + (threadIdx.x / kBX);
  const bool in = r < H && c < W;
  const float* img = a.img + size_t(t) * N;
  const float* q0 = a.p_in ? a.p_in + size_t(t) * 2 * N : nullptr;
  const float* q1 = q0 ? q0 + N : nullptr;
  auto P0 = [&](int rr, int cc) -> float {
    return (q0 && rr >= 0 && rr < H && cc >= 0 && cc < W) ? q0[size_t(rr) * W + cc] : 0.f;
  };
  auto P1 = [&](int rr, int cc) -> float {
    return (q1 && rr >= 0 && rr < H && cc >= 0 && cc < W) ? q1[size_t(rr) * W + cc] : 0.f;
  };
  // d at (rr,cc): p0[rr-1,cc]*[rr>0] - p0 + p1[rr,cc-1]*[cc>0] - p1.
  auto Dv = [&](int rr, int cc) -> float {
    return (P0(rr - 1, cc) - P0(rr, cc)) + (P1(rr, cc - 1) - P1(rr, cc));
  };

  float e = 0.f;
  if (in) {
    const size_t o = size_t(r) * W + c;
    const float d = Dv(r, c);
    const float ov = img[o] + d;
    if (a.u == a.n - 1) {
      a.out[size_t(t) * N + o] = ov;
    } else {
      const float g0 = r < H - 1 ? (img[o + W] + Dv(r + 1, c)) - ov : 0.f;
      const float g1 = c < W - 1 ? (img[o + 1] + Dv(r, c + 1)) - ov : 0.f;
      const float nrm = fast_sqrt(fmaf(g0, g0, g1 * g1));
      const float inv = fast_rcp(fmaf(a.tw, nrm, 1.f));
      float* po = a.p_out + size_t(t) * 2 * N;
      po[o] = fmaf(-0.25f, g0, P0(r, c)) * inv;
      po[N + o] = fmaf(-0.25f, g1, P1(r, c)) * inv;
      if (a.u >= 1) a.out[size_t(t) * N + o] = ov;
      e = fmaf(d, d, a.weight * nrm);
    }
  }
  if (a.u == a.n - 1) return;
  const float s = wave_sum(e);
  __shared__ float part[kBX * kBY / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int w = 0; w < kBX * kBY / 64; ++w) tot += part[w];
    atomicAdd(a.energy + size_t(t) * a.n + a.u, tot);
  }
}

// stop[t] = iteration whose image was returned: the first exit found among
// iterations [1, upto) of the frame's energy row, else `last`.
__global__ void k_stop_iter(const double* energy, int T, int stride, int upto, int last,
                            double eps, int* stop) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  const int s = stopped_at(energy + size_t(t) * stride, upto, eps);
  stop[t] = s >= 0 ? s : last;
}

}  // namespace

size_t tv_general_workspace(int64_t H, int64_t W, int64_t T, int n) {
  const size_t N = size_t(H) * W;
  return align256(2 * 2 * size_t(T) * N * 4) + align256(size_t(T) * (n > 0 ? n : 1) * 8);
}

// Per-update TV of `image` (T,H,W) into `out`; n >= 2.
int tv_general(const float* image, float* out, int64_t H, int64_t W, int64_t T, float weight,
               float eps, int n, int* stop_iter, void* ws, hipStream_t s) {
  const size_t N = size_t(H) * W;
  float* p[2];
  p[0] = static_cast<float*>(ws);
  p[1] = p[0] + 2 * size_t(T) * N;
  double* energy = reinterpret_cast<double*>(static_cast<char*>(ws) + align256(4 * size_t(T) * N * 4));
  if (int rc = fill_f64(energy, int(T) * n, 0.0, s)) return rc;
  StepArgs a{};
  a.img = image;
  a.out = out;
  a.energy = energy;
  a.H = int(H);
  a.W = int(W);
  a.T = int(T);
  a.n = n;
  a.weight = weight;
  a.tw = float(0.25 / double(weight));
  a.eps = eps;
  dim3 grid(unsigned((W + kBX - 1) / kBX), unsigned((H + kBY - 1) / kBY), unsigned(T));
  for (int u = 0; u < n; ++u) {
    a.u = u;
    a.p_in = u == 0 ? nullptr : p[u & 1];
    a.p_out = u == n - 1 ? nullptr : p[(u + 1) & 1];
    {
      ProfScope ps("tv_step", s);
      k_tv_step<<<grid, kBX * kBY, 0, s>>>(a);
    }
    SCIGAP_LAUNCH_CHECK("tv step launch");
  }
  if (stop_iter) {
    k_stop_iter<<<unsigned((T + 255) / 256), 256, 0, s>>>(energy, int(T), n, n - 1, n - 1, eps,
                                                          stop_iter);
    SCIGAP_LAUNCH_CHECK("tv stop launch");
  }
  return SCIGAP_OK;
}

}  // namespace scigap
