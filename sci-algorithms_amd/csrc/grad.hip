// Reverse-mode kernels of the unrolled, differentiable GAP-TV.
//
// Model (per sample; Phi is shared by the batch):
//   y     = sum_t Phi_t * truth_t            (or a given measurement)
//   S     = sum_t Phi_t
//   x_0   = Phi * y,  y1_0 = 0
//   for k < K:  yb = sum_t Phi_t x_k,  y1' = y1_k + (y - yb),
//               r = (y1' - yb)/S,  v = x_k + lambda_k * r * Phi,
//               x_{k+1} = TV_{w_k}(v)  (D = tv_iter_max-1 Chambolle updates,
//                                        no early exit)
// which is PnP_SCI/python/pnp_sci_algo.py:625-650 (accelerated GAP) with the
// TV step of scikit-image 0.18.3 _denoise.py:341-393 run for a fixed number
// of iterations, and per-iteration lambda_k / tv_weight_k.
//
// One Chambolle update, with G the forward difference (zero in the last row /
// column) and G^T its transpose (the "d" of _denoise.py:352-361):
//   out_u = v + G^T p_u,  g_u = G out_u,  n_u = |g_u|,  s_u = 1/(1 + c n_u),
//   a_u = p_u - tau g_u,  p_{u+1} = a_u * s_u,   c = tau / w,  tau = 1/4.
// Reverse of one update, given pbar_{u+1}:
//   abar_u = pbar_{u+1} * s_u,   sbar = <pbar_{u+1}, a_u>,
//   nbar = -c s_u^2 sbar,  cbar += -n_u s_u^2 sbar,
//   gbar_u = -tau abar_u + nbar g_u / n_u   (0 where n_u = 0),
//   obar_u = G^T gbar_u,  vbar += obar_u,  pbar_u = abar_u + G obar_u.
// The TV output out_D = v + G^T p_D starts the chain: vbar = xbar, pbar_D = G xbar.
//
// Activations are checkpointed once per outer iteration (x_k, y1' and the
// residual r_k the forward used); the TV step of one outer iteration is
// reversed by one fused tiled kernel (tv_bwd.hip) that forms v from x_k, r_k
// and Phi as the forward did and recomputes the inner dual fields in
// registers.  Layout (T,H,W) per sample.

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <vector>

#include "scigap_internal.h"

namespace scigap {

namespace {

constexpr int BX = 64, BYR = 4;  // 256-thread blocks of the per-pixel kernels
constexpr int kGroup = 4;        // frames whose loads a per-pixel kernel issues together
constexpr float kTau = 0.25f;

// Block reduction of one float into a double accumulator (scaled).
__device__ __forceinline__ void block_atomic(float v, double* dst, double scale) {
  const float s = wave_sum(v);
  __shared__ float part[BX * BYR / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0.0;
    for (int w = 0; w < BX * BYR / 64; ++w) tot += part[w];
    atomicAdd(dst, tot * scale);
  }
}

// Reverse of the projection v = x + lambda*r*Phi, r = (y1' - yb)/S,
// y1' = y1 + y - yb, yb = sum_t Phi_t x_t (per pixel, all frames), given
// vbar = dL/dv.  The per-pixel part only: rbar = lambda sum_t vbar_t Phi_t,
// lambar += sum vbar r Phi, y1bar' = y1bar_next + rbar/S, ybb = -rbar/S - y1bar',
// ybar += y1bar', Sbar = -rbar r / S.  The frame part xbar_t = vbar_t + ybb Phi_t
// is never stored: the next reverse TV step forms it on load, and the mask
// gradient (k_ub_phibar) takes ybb and Sbar from the planes written here.
__global__ void __launch_bounds__(256) k_ub_proj(
    const float* __restrict__ vbar, const float* __restrict__ phi, const float* __restrict__ S,
    const float* __restrict__ r, float lam, const float* __restrict__ y1bar_ext,
    float* __restrict__ y1bar_out, float* __restrict__ ybar, float* __restrict__ ybb_out,
    float* __restrict__ sbar_out, double* lambar, int64_t N, int T) {
  float lsum = 0.f;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < N;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float ri = r[i], si = S[i];
    double rsum = 0.0;
    float fsum = 0.f;
    // Frames in groups of kGroup with all loads of a group issued together
    // (one memory latency per group); the sums keep frame order.
    int t = 0;
    for (; t + kGroup <= T; t += kGroup) {
      float vb[kGroup], f[kGroup];
#pragma unroll
      for (int g = 0; g < kGroup; ++g) {
        vb[g] = vbar[size_t(t + g) * N + i];
        f[g] = phi[size_t(t + g) * N + i];
      }
#pragma unroll
      for (int g = 0; g < kGroup; ++g) {
        rsum = acc_prod(vb[g], f[g], rsum);
        fsum += f[g];
      }
    }
    for (; t < T; ++t) {
      const float vb = vbar[size_t(t) * N + i], f = phi[size_t(t) * N + i];
      rsum = acc_prod(vb, f, rsum);
      fsum += f;
    }
    float rbar = float(rsum);
    lsum = fmaf(rbar, ri, lsum);  // dL/dlambda = sum_t vbar*r*Phi
    rbar *= lam;
    const float y1tot = (y1bar_ext ? y1bar_ext[i] : 0.f) + rbar / si;
    y1bar_out[i] = y1tot;
    ybar[i] += y1tot;
    ybb_out[i] = -rbar / si - y1tot;
    // S = sum_t Phi_t, except where that sum is 0 and S is the constant 1.
    if (sbar_out) sbar_out[i] = fsum != 0.f ? -rbar * ri / si : 0.f;
  }
  block_atomic(lsum, lambar, 1.0);
}

// Mask gradient of one projection: phibar_t += lambda r vbar_t + ybb x_t + Sbar.
__global__ void __launch_bounds__(256) k_ub_phibar(
    const float* __restrict__ vbar, const float* __restrict__ x, const float* __restrict__ r,
    float lam, const float* __restrict__ ybb, const float* __restrict__ sbar,
    float* __restrict__ phibar, int64_t N, int T) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < N;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float lr = lam * r[i], yb = ybb[i], sb = sbar[i];
#pragma unroll kGroup
    for (int t = 0; t < T; ++t) {
      const size_t o = size_t(t) * N + i;
      phibar[o] += fmaf(lr, vbar[o], fmaf(yb, x[o], sb));
    }
  }
}

// Reverse of x_0 = Phi*y: phibar += xbar_0 * y, ybar += sum_t xbar_0 * Phi,
// with xbar_0 = vbar + ybb Phi formed here (ybb null: xbar_0 = vbar).
__global__ void __launch_bounds__(256) k_ub_init(const float* __restrict__ vbar,
                                                 const float* __restrict__ ybb,
                                                 const float* __restrict__ phi,
                                                 const float* __restrict__ y,
                                                 float* __restrict__ ybar,
                                                 float* __restrict__ phibar, int64_t N, int T) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < N;
       i += int64_t(gridDim.x) * blockDim.x) {
    double acc = 0.0;
    const float yv = y[i], yb = ybb ? ybb[i] : 0.f;
#pragma unroll kGroup
    for (int t = 0; t < T; ++t) {
      const size_t o = size_t(t) * N + i;
      const float f = phi[o];
      const float xb = ybb ? fmaf(yb, f, vbar[o]) : vbar[o];
      acc = acc_prod(xb, f, acc);
      if (phibar) phibar[o] = fmaf(xb, yv, phibar[o]);
    }
    ybar[i] += float(acc);
  }
}

__global__ void k_fill_f32(float* p, int64_t n, float v) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// dst += src (elementwise).
__global__ void k_add_f32(float* __restrict__ dst, const float* __restrict__ src, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    dst[i] += src[i];
}

unsigned grid1(int64_t n) {
  int64_t b = (n + 255) / 256;
  return unsigned(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

// phi = sigmoid(logits).
__global__ void k_sigmoid(const float* __restrict__ z, float* __restrict__ phi, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    phi[i] = 1.f / (1.f + expf(-z[i]));
}

// dL/dlogits = dL/dphi * phi * (1 - phi).
__global__ void k_sigmoid_bwd(const float* __restrict__ phibar, const float* __restrict__ phi,
                              float* __restrict__ zbar, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float f = phi[i];
    zbar[i] = phibar[i] * f * (1.f - f);
  }
}

// Reverse of y = sum_t phi_t * truth_t: phibar_t += ybar * truth_t.
__global__ void k_measure_bwd(const float* __restrict__ ybar, const float* __restrict__ truth,
                              float* __restrict__ phibar, int64_t N, int T) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < N;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float yb = ybar[i];
    for (int t = 0; t < T; ++t) phibar[size_t(t) * N + i] = fmaf(yb, truth[size_t(t) * N + i],
                                                                 phibar[size_t(t) * N + i]);
  }
}

}  // namespace

int tv_backward_tiles(int D, const float* x, const float* phi, const float* r, float lam,
                      const float* xbar, const float* ybb, float* vbar, int64_t H, int64_t W,
                      int64_t T, float w, double* wbar, hipStream_t s);

size_t unrolled_backward_workspace(int64_t H, int64_t W, int64_t T, int D) {
  (void)D;
  const size_t N = size_t(H) * W, V = N * size_t(T);
  // vbar x2, y1bar x2, ybb x2, Sbar
  return 2 * align256(V * 4) + 5 * align256(N * 4);
}

// Backward through K unrolled iterations of one sample.
int unrolled_backward(int64_t H, int64_t W, int64_t T, int K, int D, const float* lambdas,
                      const float* weights, const float* y, const float* phi, const float* S,
                      const float* ckpt_x, size_t xstride, const float* ckpt_r, size_t rstride,
                      const float* grad_x,
                      float* grad_phi, float* grad_y, double* grad_lambda, double* grad_weight,
                      void* ws, hipStream_t s) {
  const int64_t N = H * W, V = N * T;
  char* base = static_cast<char*>(ws);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* q = base + off;
    off += align256(bytes);
    return reinterpret_cast<float*>(q);
  };
  float* vbar[2] = {take(V * 4), take(V * 4)};
  float* y1bar[2] = {take(N * 4), take(N * 4)};
  float* ybb[2] = {take(N * 4), take(N * 4)};
  float* sbar = take(N * 4);

  k_fill_f32<<<grid1(N), 256, 0, s>>>(grad_y, N, 0.f);
  SCIGAP_LAUNCH_CHECK("grad fill");
  // Adjoint of x_{k+1}: grad_x for k = K-1, else vbar_{k+1} + ybb_{k+1} Phi
  // (formed on load by the reverse TV step and by k_ub_init).
  const float* vb_in = grad_x;
  const float* ybb_in = nullptr;
  const float* y1b_in = nullptr;
  int sel = 0;
  for (int k = K - 1; k >= 0; --k) {
    const float lam = lambdas[k], w = weights[k];
    const float* xk = ckpt_x + size_t(k) * xstride;
    const float* r = ckpt_r + size_t(k) * rstride;  // residual the forward used
    {
      ProfScope ps("unrolled_bwd_tv", s);
      if (int rc = tv_backward_tiles(D, xk, phi, r, lam, vb_in, ybb_in, vbar[sel], H, W, T, w,
                                     grad_weight + k, s))
        return rc;
    }
    {
      ProfScope ps("unrolled_bwd_proj", s);
      k_ub_proj<<<grid1(N), 256, 0, s>>>(vbar[sel], phi, S, r, lam, y1b_in, y1bar[sel], grad_y,
                                         ybb[sel], grad_phi ? sbar : nullptr, grad_lambda + k,
                                         N, int(T));
      SCIGAP_LAUNCH_CHECK("unrolled projection backward");
      if (grad_phi) {
        k_ub_phibar<<<grid1(N), 256, 0, s>>>(vbar[sel], xk, r, lam, ybb[sel], sbar, grad_phi,
                                             N, int(T));
        SCIGAP_LAUNCH_CHECK("unrolled mask gradient");
      }
    }
    vb_in = vbar[sel];
    ybb_in = ybb[sel];
    y1b_in = y1bar[sel];
    sel ^= 1;
  }
  // x_0 = Phi * y
  k_ub_init<<<grid1(N), 256, 0, s>>>(vb_in, ybb_in, phi, y, grad_y, grad_phi, N, int(T));
  SCIGAP_LAUNCH_CHECK("unrolled init backward");
  return SCIGAP_OK;
}

}  // namespace scigap

// ---------------------------------------------------------------------------
// C entry points
// ---------------------------------------------------------------------------

namespace scigap {

int fill_f64(double* p, int n, double v, hipStream_t s);
int unrolled_forward_iterations(int64_t H, int64_t W, int64_t T, int K, int D,
                                const float* lambdas, const float* weights, const float* y,
                                const float* phi, const float* S, float* xs, size_t xstride,
                                float* y1b, float* x_out, float* rs, size_t rstride,
                                int* nexit, float* yb_part, hipStream_t s);

namespace {

// Checkpoint of one sample: x_0..x_{K-1}, r_0..r_{K-1}, y; every plane set
// starts on a 256-byte boundary (strides in floats).
struct CkptLayout {
  size_t x = 0, r = 0, y = 0, per_sample = 0, total = 0, xstride = 0, rstride = 0;
};

CkptLayout ckpt_layout(const scigap_unrolled& u) {
  const size_t N = size_t(u.H) * u.W, V = N * size_t(u.T), K = size_t(u.iters);
  CkptLayout c;
  c.xstride = align256(V * 4) / 4;
  c.rstride = align256(N * 4) / 4;
  c.x = 0;
  c.r = K * c.xstride * 4;
  c.y = c.r + K * c.rstride * 4;
  c.per_sample = c.y + align256(N * 4);
  c.total = c.per_sample * size_t(u.batch);
  return c;
}

// Samples of a batch run concurrently on up to kLanes streams ("lanes");
// every lane has its own scratch, and lanes 1.. their own dL/dphi partial.
constexpr int kLanes = 4;

struct ULane {
  float* y1b = nullptr;     // two y1 planes (forward)
  int* nexit = nullptr;     // K
  float* yb_part = nullptr; // [min(T,8)][H*W] partial yb of frame-split launches
  float* ybar = nullptr;    // (H,W) dL/dy of one sample
  float* phibar = nullptr;  // (T,H,W) dL/dphi partial (lanes >= 1)
  void* bwd = nullptr;
};

struct UWorkspace {
  float* phi = nullptr;     // sigmoid(logits), when logits
  float* S = nullptr;       // phi_sum
  float* phibar = nullptr;  // (T,H,W) dL/dphi, when logits
  double* gl = nullptr;     // K, when the caller passes no grad_lambda
  double* gw = nullptr;     // K, when the caller passes no grad_weight
  int lanes = 1;
  ULane lane[kLanes];
  size_t bytes = 0;
};

int lanes_of(const scigap_unrolled& u) { return u.batch < kLanes ? u.batch : kLanes; }

UWorkspace carve(const scigap_unrolled& u, char* base) {
  const size_t N = size_t(u.H) * u.W, V = N * size_t(u.T), K = size_t(u.iters);
  const int D = u.tv_iter_max - 1;
  UWorkspace w;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* q = base ? base + off : nullptr;
    off += align256(bytes);
    return q;
  };
  if (u.logits) w.phi = reinterpret_cast<float*>(take(V * 4));
  w.S = reinterpret_cast<float*>(take(N * 4));
  if (u.logits) w.phibar = reinterpret_cast<float*>(take(V * 4));
  w.gl = reinterpret_cast<double*>(take(K * 8));
  w.gw = reinterpret_cast<double*>(take(K * 8));
  w.lanes = lanes_of(u);
  for (int l = 0; l < w.lanes; ++l) {
    ULane& L = w.lane[l];
    L.y1b = reinterpret_cast<float*>(take(2 * N * 4));
    L.nexit = reinterpret_cast<int*>(take(K * 4));
    L.yb_part = reinterpret_cast<float*>(take(size_t(u.T < 8 ? u.T : 8) * N * 4));
    L.ybar = reinterpret_cast<float*>(take(N * 4));
    if (l > 0) L.phibar = reinterpret_cast<float*>(take(V * 4));
    L.bwd = take(unrolled_backward_workspace(u.H, u.W, u.T, D));
  }
  w.bytes = off;
  return w;
}

// Non-blocking, default-priority streams for the lanes, created once per
// device and kept for the life of the process.
struct LanePool {
  std::mutex mu;
  std::map<int, std::vector<hipStream_t>> streams;
};
LanePool& lane_pool() {
  static LanePool p;
  return p;
}

// Fork: lanes 1.. of a call wait for everything enqueued on `s` so far.
// Join: `s` waits for every lane.  With one lane, or while `s` is being
// captured into a graph, all samples stay on `s` (no fork).
class Lanes {
 public:
  Lanes(int want, hipStream_t s) : s_(s) {
    st_[0] = s;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (want <= 1 || hipStreamIsCapturing(s, &cs) != hipSuccess ||
        cs != hipStreamCaptureStatusNone)
      return;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    LanePool& p = lane_pool();
    std::lock_guard<std::mutex> lock(p.mu);
    std::vector<hipStream_t>& v = p.streams[dev];
    while (int(v.size()) < kLanes - 1) {
      hipStream_t t = nullptr;
      if (hipStreamCreateWithFlags(&t, hipStreamNonBlocking) != hipSuccess) break;
      v.push_back(t);
    }
    const int n = std::min<int>(want, 1 + int(v.size()));
    if (n <= 1 || hipEventCreateWithFlags(&fork_, hipEventDisableTiming) != hipSuccess) return;
    if (hipEventRecord(fork_, s) != hipSuccess) return;
    for (int l = 1; l < n; ++l) {
      if (hipStreamWaitEvent(v[l - 1], fork_, 0) != hipSuccess) return;
      st_[l] = v[l - 1];
      n_ = l + 1;
    }
  }
  int count() const { return n_; }
  hipStream_t stream(int l) const { return st_[l % n_]; }
  int lane(int b) const { return b % n_; }
  // Makes `s` wait for every lane; returns a status.
  int join() {
    for (int l = 1; l < n_; ++l) {
      hipEvent_t e = nullptr;
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
        return fail(SCIGAP_EHIP, "hipEventCreate failed");
      const bool ok = hipEventRecord(e, st_[l]) == hipSuccess &&
                      hipStreamWaitEvent(s_, e, 0) == hipSuccess;
      (void)hipEventDestroy(e);
      if (!ok) return fail(SCIGAP_EHIP, "lane join failed");
    }
    n_ = 1;
    return SCIGAP_OK;
  }
  ~Lanes() {
    join();
    if (fork_) (void)hipEventDestroy(fork_);
  }

 private:
  hipStream_t s_;
  hipStream_t st_[kLanes] = {};
  hipEvent_t fork_ = nullptr;
  int n_ = 1;
};

int check_unrolled(const scigap_unrolled* u) {
  if (!u) return fail(SCIGAP_ENULL, "unrolled descriptor is NULL");
  if (int rc = check_shape(u->H, u->W, u->T)) return rc;
  if (u->H * u->W > (int64_t(1) << 30) - 8)
    return fail(SCIGAP_EINVAL, "H*W must be <= 2^30 - 8 (buffer-addressed planes)");
  if (u->batch < 1) return fail(SCIGAP_EINVAL, "batch must be >= 1");
  if (u->iters < 1) return fail(SCIGAP_EINVAL, "iters must be >= 1");
  if (u->tv_iter_max < 2 || u->tv_iter_max > 7)
    return fail(SCIGAP_EUNSUPPORTED, "tv_iter_max must be in [2, 7] (fused TV depth)");
  if (u->batch > 1 && (u->H * u->W * u->T) % 2 != 0)
    return fail(SCIGAP_EINVAL, "batch > 1 needs an even T*H*W (8-byte aligned samples)");
  return SCIGAP_OK;
}

int check_common(const scigap_unrolled* u, const float* lambdas, const float* weights,
                 const float* mask, const float* meas, const void* ckpt, size_t ckpt_bytes,
                 void* ws, size_t ws_bytes) {
  if (int rc = check_unrolled(u)) return rc;
  if (!lambdas || !weights || !mask || !meas || !ckpt)
    return fail(SCIGAP_ENULL, "unrolled: null argument");
  for (int k = 0; k < u->iters; ++k) {
    if (!(weights[k] > 0.f) || !std::isfinite(weights[k]))
      return fail(SCIGAP_EINVAL, "tv weight %d must be finite and > 0", k);
    if (!std::isfinite(lambdas[k])) return fail(SCIGAP_EINVAL, "lambda %d is not finite", k);
  }
  const size_t need = ckpt_layout(*u).total;
  if (ckpt_bytes < need)
    return fail(SCIGAP_EWORKSPACE, "checkpoint needs %zu bytes, got %zu", need, ckpt_bytes);
  const size_t wneed = carve(*u, nullptr).bytes;
  if (!ws || ws_bytes < wneed)
    return fail(SCIGAP_EWORKSPACE, "workspace needs %zu bytes, got %zu", wneed, ws_bytes);
  if ((reinterpret_cast<uintptr_t>(mask) | reinterpret_cast<uintptr_t>(ckpt)) % 8 != 0)
    return fail(SCIGAP_EINVAL, "device buffers must be 8-byte aligned");
  return SCIGAP_OK;
}

// phi (from logits when asked) and its frame sum.
int prepare_mask(const scigap_unrolled& u, const float* mask, const UWorkspace& w,
                 const float** phi, hipStream_t s) {
  const int64_t N = u.H * u.W, V = N * u.T;
  *phi = mask;
  if (u.logits) {
    k_sigmoid<<<grid1(V), 256, 0, s>>>(mask, w.phi, V);
    SCIGAP_LAUNCH_CHECK("sigmoid launch");
    *phi = w.phi;
  }
  return scigap_phi_sum(*phi, w.S, u.H, u.W, u.T, s);
}

}  // namespace
}  // namespace scigap

using namespace scigap;

extern "C" {

size_t scigap_unrolled_checkpoint_bytes(const scigap_unrolled* u) {
  if (check_unrolled(u) != SCIGAP_OK) return 0;
  return ckpt_layout(*u).total;
}

size_t scigap_unrolled_workspace_bytes(const scigap_unrolled* u) {
  if (check_unrolled(u) != SCIGAP_OK) return 0;
  return carve(*u, nullptr).bytes;
}

int scigap_unrolled_forward(const scigap_unrolled* u, const float* lambdas, const float* weights,
                            const float* mask, const float* meas, float* x_out, void* ckpt,
                            size_t ckpt_bytes, void* workspace, size_t workspace_bytes,
                            void* stream) {
  clear_error();
  if (int rc = check_common(u, lambdas, weights, mask, meas, ckpt, ckpt_bytes, workspace,
                            workspace_bytes))
    return rc;
  if (!x_out) return fail(SCIGAP_ENULL, "unrolled: null x_out");
  if (reinterpret_cast<uintptr_t>(x_out) % 8 != 0)
    return fail(SCIGAP_EINVAL, "device buffers must be 8-byte aligned");
  const int64_t N = u->H * u->W, V = N * u->T;
  const size_t outb = size_t(V) * 4 * u->batch;
  const CkptLayout cl = ckpt_layout(*u);
  if (overlaps(x_out, outb, ckpt, cl.total) || overlaps(x_out, outb, workspace, workspace_bytes))
    return fail(SCIGAP_EALIAS, "x_out overlaps the checkpoint or the workspace");
  hipStream_t s = as_stream(stream);
  const UWorkspace w = carve(*u, static_cast<char*>(workspace));
  const float* phi = nullptr;
  if (int rc = prepare_mask(*u, mask, w, &phi, s)) return rc;
  Lanes lanes(w.lanes, s);  // sample b runs on lane b % lanes
  for (int b = 0; b < u->batch; ++b) {
    const ULane& L = w.lane[lanes.lane(b)];
    hipStream_t ls = lanes.stream(b);
    char* cb = static_cast<char*>(ckpt) + cl.per_sample * b;
    float* xs = reinterpret_cast<float*>(cb + cl.x);
    float* rs = reinterpret_cast<float*>(cb + cl.r);
    float* yb = reinterpret_cast<float*>(cb + cl.y);
    if (u->measure) {
      if (int rc = scigap_forward(meas + size_t(b) * V, phi, yb, u->H, u->W, u->T, ls)) return rc;
    } else {
      hipError_t e = hipMemcpyAsync(yb, meas + size_t(b) * N, size_t(N) * 4,
                                    hipMemcpyDeviceToDevice, ls);
      if (e != hipSuccess) return hip_fail(e, "copy y");
    }
    if (int rc = scigap_adjoint(yb, phi, xs, u->H, u->W, u->T, ls)) return rc;  // x_0 = At(y)
    if (int rc = unrolled_forward_iterations(u->H, u->W, u->T, u->iters, u->tv_iter_max - 1,
                                             lambdas, weights, yb, phi, w.S, xs, cl.xstride,
                                             L.y1b, x_out + size_t(b) * V, rs, cl.rstride,
                                             L.nexit, L.yb_part, ls))
      return rc;
  }
  return lanes.join();
}

int scigap_unrolled_backward(const scigap_unrolled* u, const float* lambdas, const float* weights,
                             const float* mask, const float* meas, const void* ckpt,
                             size_t ckpt_bytes, const float* grad_x, float* grad_mask,
                             float* grad_meas, double* grad_lambda, double* grad_weight,
                             void* workspace, size_t workspace_bytes, void* stream) {
  clear_error();
  if (int rc = check_common(u, lambdas, weights, mask, meas, ckpt, ckpt_bytes, workspace,
                            workspace_bytes))
    return rc;
  if (!grad_x) return fail(SCIGAP_ENULL, "unrolled: null grad_x");
  if (u->measure && grad_meas)
    return fail(SCIGAP_EINVAL, "measure = 1: the truth videos get no gradient (grad_meas NULL)");
  const int64_t N = u->H * u->W, V = N * u->T;
  const int K = u->iters, D = u->tv_iter_max - 1;
  const CkptLayout cl = ckpt_layout(*u);
  if ((grad_mask && (overlaps(grad_mask, V * 4, workspace, workspace_bytes) ||
                     overlaps(grad_mask, V * 4, ckpt, cl.total))) ||
      (grad_meas && overlaps(grad_meas, size_t(N) * 4 * u->batch, workspace, workspace_bytes)))
    return fail(SCIGAP_EALIAS, "a gradient output overlaps the checkpoint or the workspace");
  hipStream_t s = as_stream(stream);
  const UWorkspace w = carve(*u, static_cast<char*>(workspace));
  const float* phi = nullptr;
  if (int rc = prepare_mask(*u, mask, w, &phi, s)) return rc;
  double* gl = grad_lambda ? grad_lambda : w.gl;
  double* gw = grad_weight ? grad_weight : w.gw;
  if (int rc = fill_f64(gl, K, 0.0, s)) return rc;
  if (int rc = fill_f64(gw, K, 0.0, s)) return rc;
  float* phibar = grad_mask ? (u->logits ? w.phibar : grad_mask) : nullptr;
  Lanes lanes(w.lanes, s);
  // dL/dphi: lane 0 accumulates into phibar, lanes 1.. into their own
  // partials, summed into phibar after the join.
  float* lane_phibar[kLanes] = {phibar};
  for (int l = 1; l < lanes.count(); ++l) lane_phibar[l] = phibar ? w.lane[l].phibar : nullptr;
  if (phibar) {
    for (int l = 0; l < lanes.count(); ++l) {
      k_fill_f32<<<grid1(V), 256, 0, lanes.stream(l)>>>(lane_phibar[l], V, 0.f);
      SCIGAP_LAUNCH_CHECK("gradient fill");
    }
  }
  for (int b = 0; b < u->batch; ++b) {
    const int l = lanes.lane(b);
    const ULane& L = w.lane[l];
    hipStream_t ls = lanes.stream(b);
    float* pb = lane_phibar[l];
    const char* cb = static_cast<const char*>(ckpt) + cl.per_sample * b;
    const float* xs = reinterpret_cast<const float*>(cb + cl.x);
    const float* rs = reinterpret_cast<const float*>(cb + cl.r);
    const float* yb = reinterpret_cast<const float*>(cb + cl.y);
    if (int rc = unrolled_backward(u->H, u->W, u->T, K, D, lambdas, weights, yb, phi, w.S, xs,
                                   cl.xstride, rs, cl.rstride, grad_x + size_t(b) * V, pb,
                                   L.ybar, gl, gw, L.bwd, ls))
      return rc;
    if (u->measure && pb) {
      k_measure_bwd<<<grid1(N), 256, 0, ls>>>(L.ybar, meas + size_t(b) * V, pb, N, int(u->T));
      SCIGAP_LAUNCH_CHECK("measurement backward launch");
    }
    if (!u->measure && grad_meas) {
      hipError_t e = hipMemcpyAsync(grad_meas + size_t(b) * N, L.ybar, size_t(N) * 4,
                                    hipMemcpyDeviceToDevice, ls);
      if (e != hipSuccess) return hip_fail(e, "copy dL/dy");
    }
  }
  const int nl = lanes.count();
  if (int rc = lanes.join()) return rc;
  for (int l = 1; l < nl && phibar; ++l) {
    k_add_f32<<<grid1(V), 256, 0, s>>>(phibar, lane_phibar[l], V);
    SCIGAP_LAUNCH_CHECK("gradient lane sum");
  }
  if (grad_mask && u->logits) {
    k_sigmoid_bwd<<<grid1(V), 256, 0, s>>>(w.phibar, phi, grad_mask, V);
    SCIGAP_LAUNCH_CHECK("sigmoid backward launch");
  }
  return SCIGAP_OK;
}

}  // extern "C"
