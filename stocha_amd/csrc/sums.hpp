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

// Bit pattern of an empty partial slot: a signalling NaN, which no
// arithmetic produces (NaN results are quiet), so no block sum can equal it.
constexpr uint64_t kEmptySlot = 0x7ff4deadbeef0001ull;
// Partial slots a fused reduction may use: fused_grid's at most 4096 blocks x
// kMaxAcc sums.
constexpr size_t kSumsPartials = (size_t)4096 * kMaxAcc;

// A slot's value once its writer's store is visible: relaxed agent-scope
// loads of the slot itself (coherent across the XCDs' L2s) until it no
// longer holds kEmptySlot.  The slot is the flag, so no ordering between
// different locations is needed and no cache write-back or invalidate.
// Every awaited slot has a writer that has already run or is resident, so
// the wait is microseconds; the bound (2^24 polls, ~0.5 s of sleeps) only
// guarantees that the kernel ends even if that ever failed -- the result is
// then a NaN, never a hang.
__device__ __forceinline__ double await_slot(double* p) {
  uint64_t* q = reinterpret_cast<uint64_t*>(p);
  for (uint32_t spins = 0; spins < (1u << 24); ++spins) {
    const uint64_t v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v != kEmptySlot) return __longlong_as_double((long long)v);
    __builtin_amdgcn_s_sleep(1);
  }
  return __builtin_nan("");
}

// One relaxed agent-scope load of a slot (its raw bits): issue a row of
// these together, then settle each with slot_value, so the loads overlap and
// only a slot still empty costs a wait.
__device__ __forceinline__ uint64_t load_slot(const double* p) {
  return __hip_atomic_load(reinterpret_cast<const uint64_t*>(p), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ double slot_value(double* p, uint64_t bits) {
  return bits != kEmptySlot ? __longlong_as_double((long long)bits) : await_slot(p);
}

__device__ __forceinline__ void put_slot(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void empty_slot(double* p) {
  __hip_atomic_store(reinterpret_cast<uint64_t*>(p), kEmptySlot, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// Per-path accumulation without per-thread accumulators: after each path the
// wave adds its lanes' contributions (a butterfly) into its own row of wacc
// (LDS, kFusedBlock / 64 rows of kMaxAcc sums, zeroed by the caller), so no
// FP64 sums stay live in registers across the path loop.  v[0..na): this
// lane's contributions (0 for a lane without a path).
__device__ __forceinline__ void wave_accumulate(const double* v, int na,
                                                double (*wacc)[kMaxAcc]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < kMaxAcc; ++i) {
    if (i < na) {
      const double w = wave_sum(v[i]);
      if (lane == 0) wacc[wave][i] += w;
    }
  }
}

// Block total of the wave rows (wave order): thread i < na gets sum i.
__device__ __forceinline__ double wave_rows_total(int na, double (*wacc)[kMaxAcc]) {
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x < na)
    for (int w = 0; w < kFusedBlock / 64; ++w) s += wacc[w][threadIdx.x];
  return s;
}

__device__ __forceinline__ void finish_block_sums(double s, int na, double (*red)[kMaxAcc],
                                                  double* partials,
                                                  unsigned* __restrict__ counter,
                                                  double* __restrict__ out);

// End of a fused kernel: store the block partial; the last block to arrive
// reduces all partials in a fixed order into out[0..na) and re-zeroes the
// counter (every block has incremented it by then).  partials: slots that
// hold kEmptySlot when the kernel starts (at least gridDim.x * na of them);
// each block fills its own, the last block (the counter tells which) waits
// for each slot to be filled (await_slot), adds it and empties it again for
// the next launch.  The counter only elects the last block: it orders
// nothing, so a relaxed increment suffices.
__device__ __forceinline__ void finish_sums(const double* acc, int na, double (*red)[kMaxAcc],
                                            double* partials, unsigned* __restrict__ counter,
                                            double* __restrict__ out) {
  finish_block_sums(block_total(acc, na, red), na, red, partials, counter, out);
}

// finish_sums from the block totals (thread i < na holds sum i); red is
// scratch LDS of the same shape as block_total's.
__device__ __forceinline__ void finish_block_sums(double s, int na, double (*red)[kMaxAcc],
                                                  double* partials,
                                                  unsigned* __restrict__ counter,
                                                  double* __restrict__ out) {
  if (threadIdx.x < na) put_slot(&partials[(uint64_t)blockIdx.x * na + threadIdx.x], s);
  __syncthreads();
  // the election's result goes through red[0][0] (red is free here: s is in
  // registers), not a flag of its own -- 8 more bytes of LDS would cost the
  // 23,040-byte fused kernel its seventh block per CU
  if (threadIdx.x == 0)
    red[0][0] = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1
                    ? 1.0
                    : 0.0;
  __syncthreads();
  if (red[0][0] == 0.0) return;
  double tot[kMaxAcc];
#pragma unroll
  for (int i = 0; i < kMaxAcc; ++i) tot[i] = 0.0;
  for (unsigned b = threadIdx.x; b < gridDim.x; b += kFusedBlock) {
    double* row = &partials[(uint64_t)b * na];
    uint64_t bits[kMaxAcc];
#pragma unroll
    for (int i = 0; i < kMaxAcc; ++i) bits[i] = i < na ? load_slot(row + i) : 0;
#pragma unroll
    for (int i = 0; i < kMaxAcc; ++i)
      if (i < na) tot[i] += slot_value(row + i, bits[i]);
  }
  // emptied after all reads: vmcnt counts stores too, so stores between the
  // rows would make every row's loads wait for the previous row's stores
  for (unsigned b = threadIdx.x; b < gridDim.x; b += kFusedBlock)
#pragma unroll
    for (int i = 0; i < kMaxAcc; ++i)
      if (i < na) empty_slot(&partials[(uint64_t)b * na + i]);
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
