// zr_scan.hpp -- device-wide exclusive prefix sums of u32 streams (three
// launches: tile reduce, scan of tile sums, tile scan; one launch for inputs
// of at most 16,384 entries).  The input is read through a functor so that
// values can be derived on the fly (group counts up to a device-side bound,
// first-touch tests, non-zero masses) without materialising them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zr {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

// inclusive scan within a 64-lane wavefront
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x)
{
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, (unsigned)d, 64);
    if (lane >= d) x += y;
  }
  return x;
}

// exclusive scan across a block of BLOCK threads; smem holds BLOCK/64 words
template <int BLOCK>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& total, uint32_t* smem)
{
  constexpr int NW = BLOCK / 64;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  uint32_t inc = wave_inclusive_scan(v);
  if (lane == 63) smem[wid] = inc;
  __syncthreads();
  if (wid == 0) {
    uint32_t s = (lane < NW) ? smem[lane] : 0u;
    s = wave_inclusive_scan(s);
    if (lane < NW) smem[lane] = s;
  }
  __syncthreads();
  uint32_t off = wid ? smem[wid - 1] : 0u;
  total = smem[NW - 1];
  __syncthreads();
  return off + inc - v;
}

template <class Ld>
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(Ld ld, uint32_t n, uint32_t* tile_sums)
{
  __shared__ uint32_t smem[kScanBlock / 64];
  const uint32_t base = blockIdx.x * (uint32_t)kScanTile + threadIdx.x * kScanItems;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    uint32_t i = base + k;
    if (i < n) s += ld(i);
  }
  uint32_t total;
  block_exclusive_scan<kScanBlock>(s, total, smem);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// one block of 1024 threads scans all tile sums in place; writes the grand total
__global__ __launch_bounds__(1024) void k_scan_tile_sums(uint32_t* tile_sums, uint32_t nt, uint32_t* total_out)
{
  __shared__ uint32_t smem[1024 / 64];
  const uint32_t chunk = (nt + 1023u) / 1024u;
  const uint32_t b = threadIdx.x * chunk;
  uint32_t s = 0;
  for (uint32_t k = 0; k < chunk; ++k)
    if (b + k < nt) s += tile_sums[b + k];
  uint32_t total;
  uint32_t off = block_exclusive_scan<1024>(s, total, smem);
  for (uint32_t k = 0; k < chunk; ++k)
    if (b + k < nt) {
      uint32_t t = tile_sums[b + k];
      tile_sums[b + k] = off;
      off += t;
    }
  if (threadIdx.x == 0 && total_out) *total_out = total;
}

template <class Ld>
__global__ __launch_bounds__(kScanBlock) void k_scan_down(Ld ld, uint32_t n, const uint32_t* tile_off, uint32_t* out)
{
  __shared__ uint32_t smem[kScanBlock / 64];
  const uint32_t base = blockIdx.x * (uint32_t)kScanTile + threadIdx.x * kScanItems;
  uint32_t vals[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    uint32_t i = base + k;
    vals[k] = (i < n) ? ld(i) : 0u;
    s += vals[k];
  }
  uint32_t total;
  uint32_t off = block_exclusive_scan<kScanBlock>(s, total, smem) + tile_off[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    uint32_t i = base + k;
    if (i < n) out[i] = off;
    off += vals[k];
  }
}

// inputs of at most kScanSingleMax elements: one block of 1024 threads scans
// them in one launch.  The inputs are read striped (every thread's loads
// independent and coalesced) into LDS, each thread then scans a run of
// <= 16 consecutive entries there, and the results go out striped; the
// total is written too
constexpr uint32_t kScanSingleMax = 1024u * 16u;
template <class Ld>
__global__ __launch_bounds__(1024) void k_scan_single(Ld ld, uint32_t n, uint32_t* out, uint32_t* total_out)
{
  __shared__ uint32_t buf[kScanSingleMax];
  __shared__ uint32_t smem[1024 / 64];
  const uint32_t tid = threadIdx.x;
  uint32_t v[16];
#pragma unroll
  for (uint32_t k = 0; k < 16u; ++k) {
    const uint32_t i = k * 1024u + tid;
    v[k] = i < n ? ld(i) : 0u;
  }
#pragma unroll
  for (uint32_t k = 0; k < 16u; ++k) {
    const uint32_t i = k * 1024u + tid;
    if (i < n) buf[i] = v[k];
  }
  __syncthreads();
  const uint32_t chunk = (n + 1023u) / 1024u;
  const uint32_t b = tid * chunk;
  uint32_t s = 0;
  for (uint32_t k = 0; k < chunk; ++k)
    if (b + k < n) s += buf[b + k];
  uint32_t total;
  uint32_t off = block_exclusive_scan<1024>(s, total, smem);
  for (uint32_t k = 0; k < chunk; ++k)
    if (b + k < n) {
      const uint32_t x = buf[b + k];
      buf[b + k] = off;
      off += x;
    }
  __syncthreads();
#pragma unroll
  for (uint32_t k = 0; k < 16u; ++k) {
    const uint32_t i = k * 1024u + tid;
    if (i < n) out[i] = buf[i];
  }
  if (tid == 0 && total_out) *total_out = total;
}

struct LdArray {
  const uint32_t* a;
  __device__ uint32_t operator()(uint32_t i) const { return a[i]; }
};
struct LdNonzeroMass {
  const double* m;
  __device__ uint32_t operator()(uint32_t i) const { return m[i] != 0.0 ? 1u : 0u; }
};

}  // namespace zr
