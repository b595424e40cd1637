// common.hpp -- shared host/device helpers of the gfx950 step-path library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "tdmd_hip.h"

namespace tdmd_hip {

// ---- error plumbing (thread-local last error, status codes) ----------------
void set_error(const std::string& msg);
int fail(int status, const std::string& msg);
int hip_fail(hipError_t err, const char* what, const char* file, int line);

#define TDMD_HIP_TRY(call)                                                   \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess)                                                    \
      return ::tdmd_hip::hip_fail(e_, #call, __FILE__, __LINE__);            \
  } while (0)

#define TDMD_LAUNCH_CHECK()                                                  \
  do {                                                                       \
    hipError_t e_ = hipGetLastError();                                       \
    if (e_ != hipSuccess)                                                    \
      return ::tdmd_hip::hip_fail(e_, "kernel launch", __FILE__, __LINE__);  \
  } while (0)

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

constexpr int kBlock = 256;
constexpr int kWave = 64;

inline int grid_for(int64_t n, int block = kBlock) {
  return static_cast<int>((n + block - 1) / block);
}

// ---- metal-unit constants (src/core/constants.hpp:15,28) -------------------
constexpr double kBoltzmann = 8.617333262e-5;
constexpr double kMvv2e = 1.0364269e-4;

// ---- box passed by value to kernels ---------------------------------------
struct DevBox {
  double lo[3];
  double size[3];
  int periodic[3];
};

inline DevBox make_devbox(const tdmd_box& b) {
  DevBox d;
  for (int a = 0; a < 3; ++a) {
    d.lo[a] = b.lo[a];
    d.size[a] = b.hi[a] - b.lo[a];
    d.periodic[a] = b.periodic[a] ? 1 : 0;
  }
  return d;
}

// ---- device helpers ---------------------------------------------------------

// Single minimum image, src/core/math.hpp:81-104 (one conditional +/- box).
template <class T>
__device__ __forceinline__ T min_image(T d, T bs, T half, int periodic) {
  if (periodic) {
    if (d > half)
      d -= bs;
    else if (d < -half)
      d += bs;
  }
  return d;
}

// r^2 with every product and sum rounded separately, in the CPU list's
// order (dot(), src/core/math.hpp:56-59). No fused multiply-add.
__device__ __forceinline__ double r2_unfused(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  const double xx = dx * dx;
  const double yy = dy * dy;
  const double zz = dz * dz;
  return (xx + yy) + zz;
}

// Wave-level sum of a double (wave64), result valid in every lane.
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ int wave_max_i(int v) {
  for (int off = 32; off > 0; off >>= 1) {
    int o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// Block reduction of a double into out[blockIdx.x] (fixed order: wave sums
// then a serial sum of wave partials by thread 0). blockDim multiple of 64.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double wsum[BLOCK / 64];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wsum[w] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < BLOCK / 64; ++k) t += wsum[k];
  __syncthreads();
  return t;  // valid in thread 0
}

// Fixed-order sum of `n` partials into *out (added or stored).
__global__ void sum_partials_kernel(const double* __restrict__ partials, int n,
                                    double* __restrict__ out, int accumulate);

int launch_sum_partials(const double* partials, int n, double* out,
                        int accumulate, hipStream_t stream);

}  // namespace tdmd_hip
