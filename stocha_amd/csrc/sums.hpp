// sums.hpp -- the fixed-order FP64 reductions that end the fused (terminal-
// value) kernels: butterfly over the 64-lane wavefront, LDS across the
// block's waves, one partial per block; the last block to finish sums the
// partials in a fixed order, so results do not depend on timing.
#pragma once
#include "host.hpp"

namespace stocha {

constexpr int kFusedBlock = 256;
constexpr int kMaxAcc = 16;  // accumulators per fused kernel

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block sums of acc[0..na) into red, then red's wave order (thread i < na
// gets the block total of slot i).
__device__ __forceinline__ double block_total(const double* acc, int na,
                                              double (*red)[kMaxAcc]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < kMaxAcc; ++i) {
    if (i < na) {
      const double w = wave_sum(acc[i]);
      if (lane == 0) red[wave][i] = w;
    }
  }
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x < na)
    for (int w = 0; w < kFusedBlock / 64; ++w) s += red[w][threadIdx.x];
  return s;
}

// End of a fused kernel: store the block partial; the last block to arrive
// reduces all partials in a fixed order into out[0..na) and re-zeroes the
// counter (every block has incremented it by then).  The partials are
// written and read with agent-scope relaxed atomics (sc1 loads/stores,
// coherent across the XCDs' L2s) and ordered by s_waitcnt, so no block pays
// for an L2 write-back / invalidate (__threadfence).
__device__ __forceinline__ void finish_sums(const double* acc, int na, double (*red)[kMaxAcc],
                                            double* partials, unsigned* __restrict__ counter,
                                            double* __restrict__ out) {
  __shared__ bool last;
  const double s = block_total(acc, na, red);
  if (threadIdx.x < na)
    __hip_atomic_store(&partials[(uint64_t)blockIdx.x * na + threadIdx.x], s, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  __builtin_amdgcn_s_waitcnt(0);  // the partial store has completed
  __syncthreads();
  if (threadIdx.x == 0) last = atomicAdd(counter, 1u) == gridDim.x - 1;
  __syncthreads();
  if (!last) return;
  double tot[kMaxAcc];
#pragma unroll
  for (int i = 0; i < kMaxAcc; ++i) tot[i] = 0.0;
  for (unsigned b = threadIdx.x; b < gridDim.x; b += kFusedBlock) {
#pragma unroll
    for (int i = 0; i < kMaxAcc; ++i)
      if (i < na)
        tot[i] += __hip_atomic_load(&partials[(uint64_t)b * na + i], __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();  // red is reused
  const double t = block_total(tot, na, red);
  if (threadIdx.x < na) out[threadIdx.x] = t;
  if (threadIdx.x == 0) *counter = 0u;
}

inline unsigned fused_grid(uint64_t n_paths) {
  // One path per thread up to 4096 blocks (16 resident waves on each of the
  // 256 CUs, twice over); larger runs loop.  Depends on n_paths only.
  uint64_t b = (n_paths + kFusedBlock - 1) / kFusedBlock;
  if (b > 4096) b = 4096;
  if (b == 0) b = 1;
  return (unsigned)b;
}

}  // namespace stocha
