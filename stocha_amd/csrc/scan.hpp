// scan.hpp -- exclusive prefix sum of uint64 counts on the device (the
// chunk decoders of bulk.hip place each chunk's outputs with it).
//
// Three launches, all with fixed-size blocks and a fixed combination order,
// so the result is the same on every run:
//   1. scan_reduce_kernel  each block sums its tile of kScanTile counts;
//   2. scan_bsum_kernel    one block turns the tile sums into exclusive tile
//                          offsets, kScanTile sums at a time with a carry;
//   3. scan_apply_kernel   each block scans its tile from its offset.
// A tile is kScanBlock threads x kScanItems consecutive counts; within a
// block, each thread sums its items, a wavefront scan (shuffles) and a
// scan over the block's four wave totals in LDS give the thread offsets.
#pragma once
#include "host.hpp"

namespace stocha {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr uint64_t kScanTile = (uint64_t)kScanBlock * kScanItems;

// Inclusive scan of v over the block (thread order); *total = the block sum.
__device__ __forceinline__ uint64_t block_inclusive_scan(uint64_t v, uint64_t* total) {
  __shared__ uint64_t wsum[kScanBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  if (lane == 63) wsum[wave] = v;
  __syncthreads();
  uint64_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kScanBlock / 64; ++w) {
    const uint64_t s = wsum[w];
    before += w < wave ? s : 0;
    all += s;
  }
  __syncthreads();  // wsum is reused by the next call
  *total = all;
  return v + before;
}

static __global__ __launch_bounds__(kScanBlock) void scan_reduce_kernel(
    const uint64_t* __restrict__ in, const uint64_t n, uint64_t* __restrict__ tsum) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint64_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) s += base + k < n ? in[base + k] : 0;
  uint64_t total;
  (void)block_inclusive_scan(s, &total);
  if (threadIdx.x == 0) tsum[blockIdx.x] = total;
}

// In place: tsum[0..nt) -> exclusive offsets.
static __global__ __launch_bounds__(kScanBlock) void scan_bsum_kernel(
    uint64_t* __restrict__ tsum, const uint64_t nt) {
  uint64_t carry = 0;
  for (uint64_t t0 = 0; t0 < nt; t0 += kScanTile) {
    const uint64_t base = t0 + (uint64_t)threadIdx.x * kScanItems;
    uint64_t v[kScanItems], s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      v[k] = base + k < nt ? tsum[base + k] : 0;
      s += v[k];
    }
    uint64_t total;
    uint64_t off = carry + block_inclusive_scan(s, &total) - s;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
      if (base + k < nt) tsum[base + k] = off;
      off += v[k];
    }
    carry += total;
  }
}

static __global__ __launch_bounds__(kScanBlock) void scan_apply_kernel(
    const uint64_t* __restrict__ in, const uint64_t n, const uint64_t* __restrict__ toff,
    uint64_t* __restrict__ out) {
  const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint64_t v[kScanItems], s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    s += v[k];
  }
  uint64_t total;
  uint64_t off = toff[blockIdx.x] + block_inclusive_scan(s, &total) - s;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = off;
    off += v[k];
  }
}

// Scratch words exclusive_sum_u64 needs for n counts.
inline size_t scan_tmp_words(uint64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

// out[i] = in[0] + ... + in[i-1] (out[0] = 0); tmp: scan_tmp_words(n) words.
// in and out may not overlap.  Enqueued on st.
inline int exclusive_sum_u64(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t* tmp,
                             hipStream_t st, const Err& err) {
  if (n == 0) return STOCHA_OK;
  const uint64_t nt = (n + kScanTile - 1) / kScanTile;
  if (nt > 0x7fffffffull) return err.set(STOCHA_EINVAL, "scan of %llu counts is too large",
                                         (unsigned long long)n);
  scan_reduce_kernel<<<(unsigned)nt, kScanBlock, 0, st>>>(in, n, tmp);
  scan_bsum_kernel<<<1, kScanBlock, 0, st>>>(tmp, nt);
  scan_apply_kernel<<<(unsigned)nt, kScanBlock, 0, st>>>(in, n, tmp, out);
  STOCHA_HIP(hipGetLastError());
  return STOCHA_OK;
}

}  // namespace stocha
