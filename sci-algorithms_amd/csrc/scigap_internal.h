// Internal helpers shared by the scigap HIP sources (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "scigap.h"

namespace scigap {

// Record a printf-style message for scigap_last_error() and return `code`.
int fail(int code, const char* fmt, ...);

// Record a HIP error and return SCIGAP_EHIP.
inline int hip_fail(hipError_t e, const char* what) {
  return fail(SCIGAP_EHIP, "%s: %s", what, hipGetErrorString(e));
}

// Clear the per-thread error message (called on entry of every C-ABI call).
void clear_error();

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Byte-range overlap test for device buffers.
inline bool overlaps(const void* a, size_t abytes, const void* b, size_t bbytes) {
  auto pa = reinterpret_cast<uintptr_t>(a), pb = reinterpret_cast<uintptr_t>(b);
  return a && b && pa < pb + bbytes && pb < pa + abytes;
}

// Shape check shared by every entry point: positive and small enough that
// all per-plane indices fit in 32 bits and whole volumes in 64 bits.
inline int check_shape(int64_t H, int64_t W, int64_t T) {
  if (H <= 0 || W <= 0 || T <= 0)
    return fail(SCIGAP_EINVAL, "shape must be positive, got H=%lld W=%lld T=%lld",
                (long long)H, (long long)W, (long long)T);
  if (H * W > (int64_t(1) << 31) - 1 || T > 65535)
    return fail(SCIGAP_EINVAL, "shape too large: H*W must be < 2^31 and T <= 65535");
  return SCIGAP_OK;
}

#define SCIGAP_LAUNCH_CHECK(what)                                        \
  do {                                                                   \
    hipError_t _e = hipGetLastError();                                   \
    if (_e != hipSuccess) return ::scigap::hip_fail(_e, what);           \
  } while (0)

// Records a HIP event pair around a launch while scigap_profile_begin() is
// active (prof.hip); a no-op otherwise.
bool profiling();
class ProfScope {
 public:
  ProfScope(const char* name, hipStream_t s);
  ~ProfScope();
  ProfScope(const ProfScope&) = delete;
  ProfScope& operator=(const ProfScope&) = delete;

 private:
  const char* name_;
  hipStream_t s_;
  void* a_ = nullptr;
  void* b_ = nullptr;
};

// Round `n` up to a multiple of 256 bytes (workspace carving).
inline size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

// Lane i receives lane i+1's value; lane 63 receives 0 (DPP wave_shl:1).
__device__ __forceinline__ float from_next_lane(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
      0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}
// Lane i receives lane i-1's value; lane 0 receives 0 (DPP wave_shr:1).
__device__ __forceinline__ float from_prev_lane(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(
      0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}

// Hardware square root and reciprocal (about 1 ulp; no IEEE fix-up sequence).
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Forces v to be computed at this point of the program.  Without it the
// compiler may sink a running sum into a later wave-uniform branch and keep
// every term of the sum live until then.
__device__ __forceinline__ void compute_here(float& v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void compute_here(unsigned& v) { asm volatile("" : "+v"(v)); }

// Per-pixel sums over the frames (A x = sum_t Phi_t x_t and its corrections)
// are accumulated in fp64 from the exact products of the fp32 factors and
// rounded to fp32 once.  A sequential fp32 sum loses about T ulps, which the
// residual y - A x then amplifies: at T = 1000 a solve would drift 3.6e-4 from
// the FP64 oracle, against 8e-6 with this accumulation (NumPy's pairwise fp32
// sum: 1.1e-5).  FP64 FMA issues at the FP32 rate on this GPU.
__device__ __forceinline__ double acc_prod(float a, float b, double acc) {
  return fma(double(a), double(b), acc);
}
// The same step for an fp32 accumulator (register-bound kernels, see their use).
__device__ __forceinline__ float acc_prod(float a, float b, float acc) { return fmaf(a, b, acc); }

// Sum over the 64 lanes of a wavefront as a wave-uniform value: DPP within
// each 16-lane row (every lane of a row ends with the row's sum), then the
// four row sums read from lanes 0, 16, 32, 48 (no LDS round trips).
__device__ __forceinline__ float wave_sum_uniform(float v) {
  v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0xB1, 0xf, 0xf, true));
  v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x4E, 0xf, 0xf, true));
  v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x141, 0xf, 0xf, true));
  v += __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x140, 0xf, 0xf, true));
  const int b = __builtin_bit_cast(int, v);
  const float r0 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 0));
  const float r1 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 16));
  const float r2 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 32));
  const float r3 = __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 48));
  return (r0 + r1) + (r2 + r3);
}

// Sum over the 64 lanes of a wavefront; every lane receives the total.
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

}  // namespace scigap
