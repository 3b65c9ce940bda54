// Forward model, adjoint, mask sum, layout conversion and squared-error
// reductions of the SCI forward model (PnP_SCI/python/utils.py:10-36,
// pnp_sci_algo.py:491-492).  All volumes are frame-major (T,H,W).
//
// These kernels are single-pass and HBM-bound: each reads and writes every
// byte once with 16-byte accesses along the contiguous pixel axis.

#include "scigap_internal.h"

#include <cstring>

namespace scigap {

namespace {
thread_local char g_err[512] = "";
}

int fail(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

void clear_error() { g_err[0] = 0; }

namespace {

constexpr int kThreads = 256;

// Blocks for a grid-stride loop over n items: enough to fill 256 CUs with
// several waves each, never more than the work needs.
inline unsigned grid_for(int64_t n, int per_thread = 1) {
  int64_t blocks = (n + int64_t(kThreads) * per_thread - 1) / (int64_t(kThreads) * per_thread);
  if (blocks > 256 * 16) blocks = 256 * 16;
  if (blocks < 1) blocks = 1;
  return unsigned(blocks);
}

// y = sum_t x_t * phi_t.  One thread handles 4 consecutive pixels (N % 4 == 0)
// or 1 pixel; the frame loop keeps the running sums in registers (fp64).
template <int V>
__global__ void __launch_bounds__(kThreads) k_forward(const float* __restrict__ x,
                                                      const float* __restrict__ phi,
                                                      float* __restrict__ y, int64_t N, int T) {
  const int64_t nv = N / V;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < nv;
       i += int64_t(gridDim.x) * blockDim.x) {
    if constexpr (V == 4) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      for (int t = 0; t < T; ++t) {
        float4 a = reinterpret_cast<const float4*>(x + t * N)[i];
        float4 m = reinterpret_cast<const float4*>(phi + t * N)[i];
        s0 = acc_prod(a.x, m.x, s0);
        s1 = acc_prod(a.y, m.y, s1);
        s2 = acc_prod(a.z, m.z, s2);
        s3 = acc_prod(a.w, m.w, s3);
      }
      reinterpret_cast<float4*>(y)[i] = make_float4(float(s0), float(s1), float(s2), float(s3));
    } else {
      double acc = 0.0;
      for (int t = 0; t < T; ++t) acc = acc_prod(x[t * N + i], phi[t * N + i], acc);
      y[i] = float(acc);
    }
  }
}

template <int V>
__global__ void __launch_bounds__(kThreads) k_adjoint(const float* __restrict__ y,
                                                      const float* __restrict__ phi,
                                                      float* __restrict__ x, int64_t N, int T) {
  const int64_t nv = N / V;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < nv;
       i += int64_t(gridDim.x) * blockDim.x) {
    if constexpr (V == 4) {
      float4 v = reinterpret_cast<const float4*>(y)[i];
      for (int t = 0; t < T; ++t) {
        float4 m = reinterpret_cast<const float4*>(phi + t * N)[i];
        reinterpret_cast<float4*>(x + t * N)[i] = make_float4(v.x * m.x, v.y * m.y, v.z * m.z, v.w * m.w);
      }
    } else {
      float v = y[i];
      for (int t = 0; t < T; ++t) x[t * N + i] = v * phi[t * N + i];
    }
  }
}

__global__ void __launch_bounds__(kThreads) k_phi_sum(const float* __restrict__ phi,
                                                      float* __restrict__ s, int64_t N, int T) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < N;
       i += int64_t(gridDim.x) * blockDim.x) {
    // Sequential sum over t (the reference's np.sum over a length-T axis of a
    // C-order array also adds the T values of a pixel in order for T < 8).
    float acc = 0.f;
    for (int t = 0; t < T; ++t) acc += phi[t * N + i];
    s[i] = (acc == 0.f) ? 1.f : acc;
  }
}

// Tiled transpose of a rows x cols matrix (row-major) into cols x rows.
// (H,W,T) -> (T,H,W) is rows = H*W, cols = T; the reverse swaps them.
// grid.y is capped at 65535, so blocks stride over the row tiles.
constexpr int TT = 32;
__global__ void __launch_bounds__(TT * 8) k_transpose(const float* __restrict__ src,
                                                      float* __restrict__ dst,
                                                      int64_t rows, int64_t cols) {
  __shared__ float tile[TT][TT + 1];
  const int tx = threadIdx.x % TT, ty = threadIdx.x / TT;  // 32 x 8 threads
  const int64_t c0 = int64_t(blockIdx.x) * TT;
  const int64_t row_tiles = (rows + TT - 1) / TT;
  for (int64_t by = blockIdx.y; by < row_tiles; by += gridDim.y) {
    const int64_t r0 = by * TT;
    for (int k = ty; k < TT; k += 8) {
      int64_t r = r0 + k, c = c0 + tx;
      if (r < rows && c < cols) tile[k][tx] = src[r * cols + c];
    }
    __syncthreads();
    for (int k = ty; k < TT; k += 8) {
      int64_t c = c0 + k, r = r0 + tx;  // dst row = source column
      if (r < rows && c < cols) dst[c * rows + r] = tile[tx][k];
    }
    __syncthreads();
  }
}

__global__ void k_fill_f64(double* p, int n, double v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// sum((a-b)^2) over n contiguous elements per slice; slice = blockIdx.y.
__global__ void __launch_bounds__(kThreads) k_sq_error(const float* __restrict__ a,
                                                       const float* __restrict__ b,
                                                       double* __restrict__ out, int64_t n,
                                                       int per_slice_out) {
  const int64_t base = int64_t(blockIdx.y) * n;
  // As scikit-image's mean_squared_error for float32 images: the difference
  // is squared in fp32 and the squares are accumulated in fp64.
  double acc = 0.0;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float d = a[base + i] - b[base + i];
    acc += double(d * d);
  }
  double s = wave_sum(acc);
  __shared__ double part[kThreads / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0;
    for (int w = 0; w < kThreads / 64; ++w) tot += part[w];
    atomicAdd(out + (per_slice_out ? blockIdx.y : 0), tot);
  }
}

// Per-frame SSIM sum over the interior pixels [3,H-3) x [3,W-3); fp64 like
// scikit-image (which converts both images to float64).  A 7x7 window
// centred on an interior pixel never leaves the image, so the reflect
// boundary of scipy's uniform_filter does not enter the cropped mean.
constexpr int kSB = 64, kSR = 4;
__global__ void __launch_bounds__(kSB* kSR) k_ssim(const float* __restrict__ a,
                                                  const float* __restrict__ b,
                                                  double* __restrict__ out, int H, int W,
                                                  double c1, double c2) {
  const int t = blockIdx.z;
  const size_t N = size_t(H) * W;
  const float* fa = a + t * N;
  const float* fb = b + t * N;
  const int c = 3 + blockIdx.x * kSB + (threadIdx.x % kSB);
  const int r = 3 + blockIdx.y * kSR + (threadIdx.x / kSB);
  double s = 0.0;
  if (r < H - 3 && c < W - 3) {
    double sx = 0, sy = 0, sxx = 0, syy = 0, sxy = 0;
    for (int dr = -3; dr <= 3; ++dr) {
      const size_t row = size_t(r + dr) * W;
      for (int dc = -3; dc <= 3; ++dc) {
        const double x = fa[row + c + dc], y = fb[row + c + dc];
        sx += x;
        sy += y;
        sxx += x * x;
        syy += y * y;
        sxy += x * y;
      }
    }
    const double np = 49.0, cov = np / (np - 1.0);
    const double ux = sx / np, uy = sy / np;
    const double vx = cov * (sxx / np - ux * ux);
    const double vy = cov * (syy / np - uy * uy);
    const double vxy = cov * (sxy / np - ux * uy);
    const double a1 = 2 * ux * uy + c1, a2 = 2 * vxy + c2;
    const double b1 = ux * ux + uy * uy + c1, b2 = vx + vy + c2;
    s = (a1 * a2) / (b1 * b2);
  }
  s = wave_sum(s);
  __shared__ double part[kSB * kSR / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = 0;
    for (int w = 0; w < kSB * kSR / 64; ++w) tot += part[w];
    atomicAdd(out + t, tot);
  }
}

__global__ void k_scale_f64(double* p, int n, double f) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] *= f;
}

}  // namespace

// Fill n doubles with v on `stream` (graph-capturable, unlike a memset node).
int fill_f64(double* p, int n, double v, hipStream_t s) {
  if (n <= 0) return SCIGAP_OK;
  k_fill_f64<<<(n + 255) / 256, 256, 0, s>>>(p, n, v);
  SCIGAP_LAUNCH_CHECK("fill kernel launch");
  return SCIGAP_OK;
}

}  // namespace scigap

using namespace scigap;

extern "C" {

int scigap_version(void) { return 10000; }

const char* scigap_last_error(void) { return scigap::g_err; }

int scigap_forward(const float* x, const float* phi, float* y, int64_t H, int64_t W, int64_t T,
                   void* stream) {
  clear_error();
  if (int rc = check_shape(H, W, T)) return rc;
  if (!x || !phi || !y) return fail(SCIGAP_ENULL, "scigap_forward: null pointer");
  const int64_t N = H * W;
  if (overlaps(y, N * 4, x, N * T * 4) || overlaps(y, N * 4, phi, N * T * 4))
    return fail(SCIGAP_EALIAS, "scigap_forward: y overlaps an input");
  hipStream_t s = as_stream(stream);
  bool vec = (N % 4 == 0) && ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(phi) |
                               reinterpret_cast<uintptr_t>(y)) % 16 == 0);
  if (vec)
    k_forward<4><<<grid_for(N / 4), kThreads, 0, s>>>(x, phi, y, N, int(T));
  else
    k_forward<1><<<grid_for(N), kThreads, 0, s>>>(x, phi, y, N, int(T));
  SCIGAP_LAUNCH_CHECK("scigap_forward launch");
  return SCIGAP_OK;
}

int scigap_adjoint(const float* y, const float* phi, float* x, int64_t H, int64_t W, int64_t T,
                   void* stream) {
  clear_error();
  if (int rc = check_shape(H, W, T)) return rc;
  if (!x || !phi || !y) return fail(SCIGAP_ENULL, "scigap_adjoint: null pointer");
  const int64_t N = H * W;
  if (overlaps(x, N * T * 4, y, N * 4) || overlaps(x, N * T * 4, phi, N * T * 4))
    return fail(SCIGAP_EALIAS, "scigap_adjoint: x overlaps an input");
  hipStream_t s = as_stream(stream);
  bool vec = (N % 4 == 0) && ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(phi) |
                               reinterpret_cast<uintptr_t>(y)) % 16 == 0);
  if (vec)
    k_adjoint<4><<<grid_for(N / 4), kThreads, 0, s>>>(y, phi, x, N, int(T));
  else
    k_adjoint<1><<<grid_for(N), kThreads, 0, s>>>(y, phi, x, N, int(T));
  SCIGAP_LAUNCH_CHECK("scigap_adjoint launch");
  return SCIGAP_OK;
}

int scigap_phi_sum(const float* phi, float* phi_sum, int64_t H, int64_t W, int64_t T,
                   void* stream) {
  clear_error();
  if (int rc = check_shape(H, W, T)) return rc;
  if (!phi || !phi_sum) return fail(SCIGAP_ENULL, "scigap_phi_sum: null pointer");
  const int64_t N = H * W;
  if (overlaps(phi_sum, N * 4, phi, N * T * 4))
    return fail(SCIGAP_EALIAS, "scigap_phi_sum: output overlaps phi");
  k_phi_sum<<<grid_for(N), kThreads, 0, as_stream(stream)>>>(phi, phi_sum, N, int(T));
  SCIGAP_LAUNCH_CHECK("scigap_phi_sum launch");
  return SCIGAP_OK;
}

static int transpose(const float* src, float* dst, int64_t rows, int64_t cols, void* stream,
                     const char* name) {
  if (!src || !dst) return fail(SCIGAP_ENULL, "%s: null pointer", name);
  if (overlaps(src, rows * cols * 4, dst, rows * cols * 4))
    return fail(SCIGAP_EALIAS, "%s: src and dst overlap", name);
  int64_t gx = (cols + TT - 1) / TT, gy = (rows + TT - 1) / TT;
  if (gx > (int64_t(1) << 31) - 1) return fail(SCIGAP_EINVAL, "%s: volume too large", name);
  if (gy > 65535) gy = 65535;
  k_transpose<<<dim3(unsigned(gx), unsigned(gy)), TT * 8, 0, as_stream(stream)>>>(src, dst, rows,
                                                                                 cols);
  SCIGAP_LAUNCH_CHECK(name);
  return SCIGAP_OK;
}

int scigap_hwt_to_thw(const float* src, float* dst, int64_t H, int64_t W, int64_t T,
                      void* stream) {
  clear_error();
  if (int rc = check_shape(H, W, T)) return rc;
  if (T == 1) {
    if (!src || !dst) return fail(SCIGAP_ENULL, "scigap_hwt_to_thw: null pointer");
    hipError_t e = hipMemcpyAsync(dst, src, H * W * 4, hipMemcpyDeviceToDevice, as_stream(stream));
    return e == hipSuccess ? SCIGAP_OK : hip_fail(e, "scigap_hwt_to_thw copy");
  }
  return transpose(src, dst, H * W, T, stream, "scigap_hwt_to_thw");
}

int scigap_thw_to_hwt(const float* src, float* dst, int64_t H, int64_t W, int64_t T,
                      void* stream) {
  clear_error();
  if (int rc = check_shape(H, W, T)) return rc;
  if (T == 1) {
    if (!src || !dst) return fail(SCIGAP_ENULL, "scigap_thw_to_hwt: null pointer");
    hipError_t e = hipMemcpyAsync(dst, src, H * W * 4, hipMemcpyDeviceToDevice, as_stream(stream));
    return e == hipSuccess ? SCIGAP_OK : hip_fail(e, "scigap_thw_to_hwt copy");
  }
  return transpose(src, dst, T, H * W, stream, "scigap_thw_to_hwt");
}

int scigap_sq_error(const float* a, const float* b, double* out, int64_t H, int64_t W, int64_t T,
                    int per_frame, void* stream) {
  clear_error();
  if (int rc = check_shape(H, W, T)) return rc;
  if (!a || !b || !out) return fail(SCIGAP_ENULL, "scigap_sq_error: null pointer");
  hipStream_t s = as_stream(stream);
  const int64_t N = H * W;
  if (int rc = fill_f64(out, per_frame ? int(T) : 1, 0.0, s)) return rc;
  if (per_frame) {
    unsigned gx = grid_for(N) / unsigned(T) + 1;
    if (gx > 1024) gx = 1024;
    k_sq_error<<<dim3(gx, unsigned(T)), kThreads, 0, s>>>(a, b, out, N, 1);
  } else {
    k_sq_error<<<dim3(grid_for(N * T), 1), kThreads, 0, s>>>(a, b, out, N * T, 0);
  }
  SCIGAP_LAUNCH_CHECK("scigap_sq_error launch");
  return SCIGAP_OK;
}

int scigap_ssim(const float* a, const float* b, double* out, int64_t H, int64_t W, int64_t T,
                float data_range, void* stream) {
  clear_error();
  if (int rc = check_shape(H, W, T)) return rc;
  if (!a || !b || !out) return fail(SCIGAP_ENULL, "scigap_ssim: null pointer");
  if (H < 7 || W < 7) return fail(SCIGAP_EINVAL, "win_size exceeds image extent (H, W >= 7 needed)");
  hipStream_t s = as_stream(stream);
  if (int rc = fill_f64(out, int(T), 0.0, s)) return rc;
  const double R = data_range, c1 = (0.01 * R) * (0.01 * R), c2 = (0.03 * R) * (0.03 * R);
  dim3 grid(unsigned((W - 6 + kSB - 1) / kSB), unsigned((H - 6 + kSR - 1) / kSR), unsigned(T));
  k_ssim<<<grid, kSB * kSR, 0, s>>>(a, b, out, int(H), int(W), c1, c2);
  SCIGAP_LAUNCH_CHECK("scigap_ssim launch");
  k_scale_f64<<<unsigned((T + 255) / 256), 256, 0, s>>>(out, int(T), 1.0 / (double(H - 6) * double(W - 6)));
  SCIGAP_LAUNCH_CHECK("scigap_ssim scale");
  return SCIGAP_OK;
}

}  // extern "C"
