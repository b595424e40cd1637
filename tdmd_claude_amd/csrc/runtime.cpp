// runtime.cpp -- memory, streams, events and error plumbing of the C-ABI
// (DeviceBuffer<T>, src/core/device_buffer.cuh:29-116, re-expressed as
// status-returning calls).
#include <cstring>
#include <string>

#include "common.hpp"

namespace tdmd_hip {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int status, const std::string& msg) {
  g_last_error = msg;
  return status;
}

int hip_fail(hipError_t err, const char* what, const char* file, int line) {
  g_last_error = std::string("HIP error ") + hipGetErrorName(err) + " (" +
                 hipGetErrorString(err) + ") in " + what + " at " + file +
                 ":" + std::to_string(line);
  return TDMD_E_HIP;
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v,
                                size_t n) {
  size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void copy_u32_kernel(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                                size_t n) {
  size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

__global__ void sum_partials_kernel(const double* __restrict__ partials, int n,
                                    double* __restrict__ out, int accumulate) {
  // One block of 256 threads: strided per-thread sums in fixed order, then
  // a fixed-order combination. Same inputs -> same bits.
  __shared__ double s[256];
  double t = 0.0;
  for (int k = threadIdx.x; k < n; k += 256) t += partials[k];
  s[threadIdx.x] = t;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (static_cast<int>(threadIdx.x) < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = accumulate ? (*out + s[0]) : s[0];
}

int launch_sum_partials(const double* partials, int n, double* out,
                        int accumulate, hipStream_t stream) {
  sum_partials_kernel<<<1, 256, 0, stream>>>(partials, n, out, accumulate);
  TDMD_LAUNCH_CHECK();
  return TDMD_OK;
}

}  // namespace tdmd_hip

using namespace tdmd_hip;

extern "C" {

int tdmd_abi_version(void) { return TDMD_ABI_VERSION; }

const char* tdmd_last_error(void) { return g_last_error.c_str(); }

int tdmd_device_count(int32_t* count) {
  int c = 0;
  TDMD_HIP_TRY(hipGetDeviceCount(&c));
  *count = c;
  return TDMD_OK;
}

int tdmd_set_device(int32_t device) {
  TDMD_HIP_TRY(hipSetDevice(device));
  return TDMD_OK;
}

int tdmd_device_name(char* buf, int32_t buflen) {
  int dev = 0;
  TDMD_HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t p;
  TDMD_HIP_TRY(hipGetDeviceProperties(&p, dev));
  std::snprintf(buf, static_cast<size_t>(buflen), "%s (%s, %d CUs)", p.name,
                p.gcnArchName, p.multiProcessorCount);
  return TDMD_OK;
}

int tdmd_malloc(void** dptr, size_t bytes) {
  if (!dptr) return fail(TDMD_E_INVALID, "tdmd_malloc: null out pointer");
  *dptr = nullptr;
  if (bytes == 0) return TDMD_OK;
  hipError_t e = hipMalloc(dptr, bytes);
  if (e == hipErrorOutOfMemory)
    return fail(TDMD_E_NOMEM, "tdmd_malloc: out of device memory (" +
                                  std::to_string(bytes) + " bytes)");
  if (e != hipSuccess) return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
  return TDMD_OK;
}

int tdmd_free(void* dptr) {
  if (dptr) TDMD_HIP_TRY(hipFree(dptr));
  return TDMD_OK;
}

int tdmd_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return TDMD_OK;
  if (stream) {
    TDMD_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
  } else {
    TDMD_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  }
  return TDMD_OK;
}

int tdmd_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return TDMD_OK;
  if (stream) {
    TDMD_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
  } else {
    TDMD_HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  }
  return TDMD_OK;
}

int tdmd_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return TDMD_OK;
  TDMD_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
  if (!stream) TDMD_HIP_TRY(hipStreamSynchronize(nullptr));
  return TDMD_OK;
}

int tdmd_fill_u32(void* dst, uint32_t value, size_t count, void* stream) {
  if (count == 0) return TDMD_OK;
  if (!dst) return fail(TDMD_E_INVALID, "tdmd_fill_u32: null pointer");
  int grid = static_cast<int>((count + kBlock - 1) / kBlock);
  if (grid > 4096) grid = 4096;
  fill_u32_kernel<<<grid, kBlock, 0, as_stream(stream)>>>(static_cast<uint32_t*>(dst), value, count);
  TDMD_LAUNCH_CHECK();
  return TDMD_OK;
}

int tdmd_copy_u32(void* dst, const void* src, size_t count, void* stream) {
  if (count == 0) return TDMD_OK;
  if (!dst || !src) return fail(TDMD_E_INVALID, "tdmd_copy_u32: null pointer");
  copy_u32_kernel<<<2048, kBlock, 0, as_stream(stream)>>>(
      static_cast<uint32_t*>(dst), static_cast<const uint32_t*>(src), count);
  TDMD_LAUNCH_CHECK();
  return TDMD_OK;
}

int tdmd_stream_create(void** stream) {
  hipStream_t s;
  TDMD_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return TDMD_OK;
}

int tdmd_stream_destroy(void* stream) {
  if (stream) TDMD_HIP_TRY(hipStreamDestroy(as_stream(stream)));
  return TDMD_OK;
}

int tdmd_stream_synchronize(void* stream) {
  TDMD_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
  return TDMD_OK;
}

int tdmd_device_synchronize(void) {
  TDMD_HIP_TRY(hipDeviceSynchronize());
  return TDMD_OK;
}

int tdmd_event_create(void** ev) {
  hipEvent_t e;
  TDMD_HIP_TRY(hipEventCreate(&e));
  *ev = e;
  return TDMD_OK;
}

int tdmd_event_destroy(void* ev) {
  if (ev) TDMD_HIP_TRY(hipEventDestroy(static_cast<hipEvent_t>(ev)));
  return TDMD_OK;
}

int tdmd_event_record(void* ev, void* stream) {
  TDMD_HIP_TRY(hipEventRecord(static_cast<hipEvent_t>(ev), as_stream(stream)));
  return TDMD_OK;
}

int tdmd_event_elapsed_ms(void* start, void* stop, float* ms) {
  TDMD_HIP_TRY(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
  TDMD_HIP_TRY(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start),
                                   static_cast<hipEvent_t>(stop)));
  return TDMD_OK;
}

}  // extern "C"
