// scan.cpp -- atomics-free exclusive prefix scan and max of int32 counts.
//
// Replaces the reference's cub::DeviceScan::ExclusiveSum +
// cub::DeviceReduce::Max + pack_meta_kernel chain
// (src/neighbors/device_neighbor_list.cu:147-154,211-243). Three launches:
//   1. tile_reduce: each 256-thread block reduces a 4096-element tile to
//      (sum, max) -- wave shuffles + LDS, no atomics;
//   2. tile_scan: one block turns the tile sums into exclusive tile offsets
//      and writes meta = {total, max};
//   3. tile_apply: each block rescans its tile and adds its offset.
// Integer arithmetic, so the result is exact and order-independent.
#include "common.hpp"

namespace tdmd_hip {

namespace {

constexpr int kScanBlock = 256;
constexpr int kItems = 16;
constexpr int kTile = kScanBlock * kItems;  // 4096

__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
  for (int off = 1; off < 64; off <<= 1) {
    int o = __shfl_up(v, off, 64);
    if (lane >= off) v += o;
  }
  return v;
}

// Exclusive scan of one value per thread across a 256-thread block.
// Returns the exclusive prefix; *total gets the block total.
__device__ __forceinline__ int block_excl_scan(int v, int* total) {
  __shared__ int wtot[kScanBlock / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int incl = wave_incl_scan(v, lane);
  if (lane == 63) wtot[w] = incl;
  __syncthreads();
  int base = 0, tot = 0;
  for (int k = 0; k < kScanBlock / 64; ++k) {
    if (k < w) base += wtot[k];
    tot += wtot[k];
  }
  __syncthreads();
  *total = tot;
  return base + incl - v;
}

__global__ void __launch_bounds__(kScanBlock)
tile_reduce_kernel(const int32_t* __restrict__ in, int64_t n,
                   int32_t* __restrict__ tile_sum, int32_t* __restrict__ tile_max) {
  __shared__ int ssum[kScanBlock / 64], smax[kScanBlock / 64];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile;
  int s = 0, m = 0;
  for (int k = 0; k < kItems; ++k) {
    const int64_t i = base + static_cast<int64_t>(k) * kScanBlock + threadIdx.x;
    if (i < n) {
      const int v = in[i];
      s += v;
      m = v > m ? v : m;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off, 64);
    const int o = __shfl_xor(m, off, 64);
    m = o > m ? o : m;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    ssum[w] = s;
    smax[w] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int ts = 0, tm = 0;
    for (int k = 0; k < kScanBlock / 64; ++k) {
      ts += ssum[k];
      tm = smax[k] > tm ? smax[k] : tm;
    }
    tile_sum[blockIdx.x] = ts;
    tile_max[blockIdx.x] = tm;
  }
}

__global__ void __launch_bounds__(kScanBlock)
tile_scan_kernel(int32_t* __restrict__ tile_sum, const int32_t* __restrict__ tile_max,
                 int ntiles, int32_t* __restrict__ meta) {
  int carry = 0, mx = 0;
  for (int start = 0; start < ntiles; start += kScanBlock) {
    const int t = start + static_cast<int>(threadIdx.x);
    const int v = t < ntiles ? tile_sum[t] : 0;
    const int m = t < ntiles ? tile_max[t] : 0;
    mx = m > mx ? m : mx;
    int chunk_total;
    const int ex = block_excl_scan(v, &chunk_total);
    if (t < ntiles) tile_sum[t] = carry + ex;
    carry += chunk_total;
  }
  // max across threads
  __shared__ int smax[kScanBlock];
  smax[threadIdx.x] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = 0;
    for (int k = 0; k < kScanBlock; ++k) m = smax[k] > m ? smax[k] : m;
    if (meta) {
      meta[0] = carry;
      meta[1] = m;
    }
  }
}

__global__ void __launch_bounds__(kScanBlock)
tile_apply_kernel(const int32_t* __restrict__ in, int32_t* __restrict__ out, int64_t n,
                  const int32_t* __restrict__ tile_off) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kTile;
  // Each thread owns kItems consecutive elements: local serial scan, then a
  // block scan of the thread totals.
  int v[kItems];
  int tsum = 0;
  for (int k = 0; k < kItems; ++k) {
    const int64_t i = base + static_cast<int64_t>(threadIdx.x) * kItems + k;
    v[k] = i < n ? in[i] : 0;
    tsum += v[k];
  }
  int total;
  int run = block_excl_scan(tsum, &total) + tile_off[blockIdx.x];
  for (int k = 0; k < kItems; ++k) {
    const int64_t i = base + static_cast<int64_t>(threadIdx.x) * kItems + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
}

}  // namespace

size_t scan_workspace_bytes(int64_t n) {
  const int64_t ntiles = (n + kTile - 1) / kTile;
  return static_cast<size_t>(2 * (ntiles > 0 ? ntiles : 1)) * sizeof(int32_t) + 256;
}

// The tile_reduce pass reads elements in a strided pattern and tile_apply in
// a blocked one; both cover the same kTile range per block, so the tile
// offsets line up.
int scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, int32_t* meta,
                       void* work, hipStream_t stream) {
  if (n <= 0) {
    if (meta) {
      // Graph-safe zero of meta via a tiny fill.
      return tdmd_fill_u32(meta, 0u, 2, stream);
    }
    return TDMD_OK;
  }
  const int64_t ntiles64 = (n + kTile - 1) / kTile;
  if (ntiles64 > (1 << 30)) return fail(TDMD_E_INVALID, "scan: n too large");
  const int ntiles = static_cast<int>(ntiles64);
  int32_t* tile_sum = static_cast<int32_t*>(work);
  int32_t* tile_max = tile_sum + ntiles;
  tile_reduce_kernel<<<ntiles, kScanBlock, 0, stream>>>(in, n, tile_sum, tile_max);
  TDMD_LAUNCH_CHECK();
  tile_scan_kernel<<<1, kScanBlock, 0, stream>>>(tile_sum, tile_max, ntiles, meta);
  TDMD_LAUNCH_CHECK();
  tile_apply_kernel<<<ntiles, kScanBlock, 0, stream>>>(in, out, n, tile_sum);
  TDMD_LAUNCH_CHECK();
  return TDMD_OK;
}

}  // namespace tdmd_hip

using namespace tdmd_hip;

extern "C" {

int tdmd_scan_workspace_bytes(int64_t n, size_t* bytes) {
  if (!bytes || n < 0) return fail(TDMD_E_INVALID, "scan_workspace_bytes: bad args");
  *bytes = scan_workspace_bytes(n);
  return TDMD_OK;
}

int tdmd_scan_exclusive_i32(const int32_t* d_in, int32_t* d_out, int64_t n,
                            int32_t* d_meta_out, void* d_work, void* stream) {
  if (n > 0 && (!d_in || !d_out || !d_work))
    return fail(TDMD_E_INVALID, "scan: null pointer");
  return scan_exclusive_i32(d_in, d_out, n, d_meta_out, d_work, as_stream(stream));
}

}  // extern "C"
