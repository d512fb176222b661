// lsmc.hip -- Longstaff-Schwartz American option (lsmc.rs:25-123) on the
// device.
//
// Layout: the GBM paths (mu = r, q = 0, block-split streams as gbm.rs) are
// written date-major, P[j][i] = S_{j+1} of path i, so every per-date pass
// reads one contiguous column with coalesced loads.  Cash flows cf[i] stay
// FP64.  Per exercise date d = steps-1 .. 1 (lsmc.rs:50-99):
//   pass A   exercise at d+1 (with the coefficients solved there), discount,
//            ITM count and shifted sums of (S-K), (S-K)^2   -> block partials
//   reduce A one block: ITM count, mean, population std (floor 1e-12), skip
//            flag when count < ncols+1 (lsmc.rs:62-64,67-71)
//   pass B   FP64 moments of x = (S-mean)/std over ITM paths: sum x^k for
//            k = 0..2*deg and sum cf*x^k for k = 0..deg      -> block partials
//   solve    one block: reduce, form the (deg+1)^2 normal equations, solve in
//            FP64 by Cholesky; this replaces faer's QR least squares
//            (lsmc.rs:83-84) -- same minimiser, different rounding.
// After date 1: exercise at 1, final discount, mean and two-pass variance
// (lsmc.rs:102-112).  All reductions are fixed-order.
#include <stdlib.h>

#include "host.hpp"
#include "drivers.hpp"
#include "sums.hpp"

namespace stocha {

constexpr int kLsBlock = 256;
constexpr int kMaxCols = 5;                      // poly_degree <= 4
constexpr int kMaxMom = 2 * (kMaxCols - 1) + 1;  // sum x^0..x^8
constexpr int kMaxAccB = kMaxMom + kMaxCols;

struct DateState {
  double mean, std;
  double coeff[kMaxCols];
  uint64_t itm;
  int skip;  // 1: no regression / exercise at this date
};

__device__ __forceinline__ double intrinsic(double s, double k, int is_put) {
  double v = is_put ? k - s : s - k;
  return v > 0.0 ? v : 0.0;
}

// compiler-rt __powidf2, the lowering of f64::powi (lsmc.rs:78,94)
__device__ __forceinline__ double powi_rt(double a, int b) {
  double r = 1.0;
  for (;;) {
    if (b & 1) r *= a;
    b /= 2;
    if (b == 0) break;
    a *= a;
  }
  return r;
}

// ------------------------------------------------------------ generation --
// GBM paths on the queue schedule (drivers.hpp); the step hook writes S_{j+1}
// of path i to P[j * n + i] (coalesced across the wave).  Lanes past n step
// path 0's stream and write nothing.  FP32: S = s0 * 2^(ls * log2 e) in FP32.
template <bool kFp32>
struct DateMajorHook {
  void* paths;
  uint64_t n, i;
  bool valid;
  float s0f;
  template <class M>
  __device__ __forceinline__ void operator()(uint32_t j, const M& m) const {
    if (!valid) return;
    if constexpr (kFp32)
      ((float*)paths)[(uint64_t)j * n + i] = s0f * __builtin_amdgcn_exp2f(m.ls * 1.44269504088896341f);
    else
      ((double*)paths)[(uint64_t)j * n + i] = m.s;
  }
};

// The same stores through a buffer resource per date column (columns of at
// most 2^31 - 1 bytes): the column base is wave-uniform (scalar registers)
// and the lane's byte offset is fixed for the path, so a step's store takes
// no per-lane address arithmetic; lanes past n have offsets outside the
// record range and store nothing (no branch around the store).
template <bool kFp32>
struct DateMajorBufHook {
  using T = typename std::conditional<kFp32, float, double>::type;
  const T* paths;
  uint32_t n_bytes;  // one column
  uint32_t off;      // this lane's byte offset in a column
  uint64_t n;
  float s0f;
  template <class M>
  __device__ __forceinline__ void operator()(uint32_t j, const M& m) const {
    const __amdgpu_buffer_rsrc_t col = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(paths + (uint64_t)j * n), 0, (int)n_bytes, 0x00020000);
    if constexpr (kFp32)
      __builtin_amdgcn_raw_buffer_store_b32(
          __builtin_bit_cast(uint32_t, s0f * __builtin_amdgcn_exp2f(m.ls * 1.44269504088896341f)),
          col, (int)off, 0, 0);
    else
      __builtin_amdgcn_raw_buffer_store_b64(
          __builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned int, m.s), col,
          (int)off, 0, 0);
  }
};

// Block size of the path kernel.  (512-thread FP32 blocks, 8 waves per SIMD
// and 1,048,576 paths in exactly two rounds of resident threads, measured
// slower: 168 against 147 us at 1M x 50.)
template <bool kFp32>
constexpr int gen_block() {
  return kLsBlock;
}

template <bool kFp32, bool kBufStore>
__global__ __launch_bounds__(gen_block<kFp32>()) void lsmc_gen_kernel(
    const GbmConst mc, const StreamDesc sd, const uint64_t n, const uint32_t steps,
    void* __restrict__ paths, double* __restrict__ cf, const double k, const int is_put) {
  using Src = GbmSrc<kFp32>;
  using Model = typename std::conditional<kFp32, Gbm32, Gbm64>::type;
  constexpr uint32_t kSlots = kFp32 ? 16 : 8;
  constexpr int kB = gen_block<kFp32>();
  __shared__ typename Src::Tables zig;
  __shared__ typename Src::Rec q[kSlots][kB];
  zig_load(&zig);
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * kB;
  // the first state of this thread's next path: stream_at once, then one
  // affine step of `stride` streams per path (50-step paths would otherwise
  // spend ~15% of their work finding their streams)
  const Jump next = stream_stride(&sd, stride);
  U128 first = stream_at(&sd, (uint64_t)blockIdx.x * kB + threadIdx.x);
  for (uint64_t base = (uint64_t)blockIdx.x * kB; base < n; base += stride) {
    const uint64_t i = base + threadIdx.x;
    typename std::conditional<kBufStore, DateMajorBufHook<kFp32>, DateMajorHook<kFp32>>::type hook;
    hook.paths = static_cast<decltype(hook.paths)>(paths);
    hook.n = n;
    hook.s0f = (float)mc.s0;
    if constexpr (kBufStore) {
      using T = typename DateMajorBufHook<kFp32>::T;
      hook.n_bytes = (uint32_t)(n * sizeof(T));
      // past n: an offset at or beyond the column's bytes
      hook.off = i < n ? (uint32_t)(i * sizeof(T)) : hook.n_bytes;
    } else {
      hook.i = i;
      hook.valid = i < n;
    }
    Pcg g{first};
    first = add128(mul128(next.a, first), next.c);
    (void)queue_path<Src, Model, kSlots, kB>(mc, g, sd.inc, steps, &zig, q, false, hook);
    if (cf != nullptr && i < n) {
      double last;
      if constexpr (kFp32)
        last = (double)((const float*)paths)[(uint64_t)(steps - 1) * n + i];
      else
        last = ((const double*)paths)[(uint64_t)(steps - 1) * n + i];
      cf[i] = intrinsic(last, k, is_put);  // lsmc.rs:45-47
    }
  }
}

template <bool kFp32>
__device__ __forceinline__ double spot_at(const void* paths, uint64_t n, uint64_t date, uint64_t i) {
  // date in 1..steps -> column date-1
  if constexpr (kFp32)
    return (double)((const float*)paths)[(date - 1) * n + i];
  else
    return ((const double*)paths)[(date - 1) * n + i];
}

// One date's column of the path store (n values) as a buffer resource: the
// column base sits in SGPRs and each load takes a 32-bit byte offset, so the
// per-path addresses of a thread cost no 64-bit registers; offsets at or past
// n values read 0 (the hardware range check on num_records).
template <bool kFp32>
struct SpotCol {
  using T = typename std::conditional<kFp32, float, double>::type;
  __amdgpu_buffer_rsrc_t r;
  __device__ __forceinline__ SpotCol(const void* paths, uint64_t n, uint64_t date)
      : r(__builtin_amdgcn_make_buffer_rsrc((void*)((const T*)paths + (date - 1) * n), 0,
                                            (int)(n * sizeof(T)), 0x00020000)) {}
  // value at byte offset `off` of the column
  __device__ __forceinline__ T at(uint32_t off) const {
    if constexpr (kFp32)
      return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    else
      return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
  }
  __device__ __forceinline__ T operator[](uint32_t i) const { return at(i * (uint32_t)sizeof(T)); }
};

// A thread's byte offset into a column, made opaque once per date so that the
// per-path offsets (this plus q x the grid's stride) are formed at each load
// instead of being kept live, 16 of them, across the whole induction.
__device__ __forceinline__ uint32_t fresh_offset(uint32_t t0, uint32_t esz) {
  uint32_t off = t0 * esz;
  asm volatile("" : "+v"(off));
  return off;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int NA>
__device__ __forceinline__ void block_store(double (&acc)[NA], int na, double* __restrict__ part) {
  __shared__ double red[kLsBlock / 64][kMaxAccB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < NA; ++a) {
    if (a < na) {
      double w = wave_sum_d(acc[a]);
      if (lane == 0) red[wave][a] = w;
    }
  }
  __syncthreads();
  if (threadIdx.x < na) {
    double s = 0.0;
    for (int w = 0; w < kLsBlock / 64; ++w) s += red[w][threadIdx.x];
    part[(uint64_t)blockIdx.x * na + threadIdx.x] = s;
  }
}

// Exercise at date `ex_date` (if it had a regression), then discount and
// gather ITM statistics at date `date` (0 = final pass: mean of cf).
template <bool kFp32>
__global__ __launch_bounds__(kLsBlock) void lsmc_pass_a(const void* __restrict__ paths,
                                                        double* __restrict__ cf, const uint64_t n,
                                                        const uint64_t ex_date,
                                                        const DateState* __restrict__ ds,
                                                        const uint64_t date, const double k,
                                                        const int is_put, const int ncols,
                                                        const double discount,
                                                        double* __restrict__ part) {
  double acc[3] = {0.0, 0.0, 0.0};
  const DateState* ex = ex_date ? &ds[ex_date] : nullptr;
  const bool do_ex = ex && !ex->skip;
  double mean = 0.0, sd = 1.0, c[kMaxCols] = {0, 0, 0, 0, 0};
  if (do_ex) {
    mean = ex->mean;
    sd = ex->std;
    for (int j = 0; j < ncols; ++j) c[j] = ex->coeff[j];
  }
  const uint64_t stride = (uint64_t)gridDim.x * kLsBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kLsBlock + threadIdx.x; i < n; i += stride) {
    double v = cf[i];
    if (do_ex) {  // lsmc.rs:91-98
      const double sp = spot_at<kFp32>(paths, n, ex_date, i);
      const double exv = intrinsic(sp, k, is_put);
      if (exv > 0.0) {
        const double x = (sp - mean) / sd;
        double cont = 0.0;
        for (int j = 0; j < ncols; ++j) cont += c[j] * powi_rt(x, j);
        if (exv > (cont > 0.0 ? cont : 0.0)) v = exv;
      }
    }
    v *= discount;  // lsmc.rs:52-54 (or :102-104 on the final pass)
    cf[i] = v;
    if (date) {
      const double sp = spot_at<kFp32>(paths, n, date, i);
      if (intrinsic(sp, k, is_put) > 0.0) {
        const double u = sp - k;
        acc[0] += 1.0;
        acc[1] += u;
        acc[2] += u * u;
      }
    } else {
      acc[0] += v;
    }
  }
  block_store<3>(acc, date ? 3 : 1, part);
}

// Fixed-order reduction of n_blocks x na partials into out[0..na) (one block).
__device__ void reduce_block_partials(const double* __restrict__ part, int n_blocks, int na,
                                      double* out) {
  __shared__ double buf[kLsBlock];
  for (int a = 0; a < na; ++a) {
    double s = 0.0;
    for (int b = threadIdx.x; b < n_blocks; b += kLsBlock) s += part[(uint64_t)b * na + a];
    buf[threadIdx.x] = s;
    __syncthreads();
    for (int o = kLsBlock / 2; o > 0; o >>= 1) {
      if (threadIdx.x < o) buf[threadIdx.x] += buf[threadIdx.x + o];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[a] = buf[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kLsBlock) void lsmc_reduce_a(const double* __restrict__ part,
                                                          const int n_blocks, const uint64_t date,
                                                          const double k, const int ncols,
                                                          DateState* __restrict__ ds) {
  __shared__ double tot[3];
  reduce_block_partials(part, n_blocks, 3, tot);
  if (threadIdx.x == 0) {
    DateState& d = ds[date];
    const double m = tot[0];
    d.itm = (uint64_t)m;
    if (d.itm < (uint64_t)ncols + 1) {
      d.skip = 1;
      d.mean = 0.0;
      d.std = 1.0;
    } else {
      const double mu = tot[1] / m;  // mean of (S - K)
      double var = tot[2] / m - mu * mu;
      var = var > 0.0 ? var : 0.0;
      double sdv = sqrt(var);
      d.mean = mu + k;
      d.std = sdv > 1e-12 ? sdv : 1e-12;
      d.skip = 0;
    }
  }
}

template <bool kFp32>
__global__ __launch_bounds__(kLsBlock) void lsmc_pass_b(const void* __restrict__ paths,
                                                        const double* __restrict__ cf,
                                                        const uint64_t n, const uint64_t date,
                                                        const DateState* __restrict__ ds,
                                                        const double k, const int is_put,
                                                        const int ncols,
                                                        double* __restrict__ part) {
  const DateState& d = ds[date];
  if (d.skip) return;
  const int nm = 2 * (ncols - 1) + 1;
  double acc[kMaxAccB];
#pragma unroll
  for (int a = 0; a < kMaxAccB; ++a) acc[a] = 0.0;
  const double mean = d.mean, sd = d.std;
  const uint64_t stride = (uint64_t)gridDim.x * kLsBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kLsBlock + threadIdx.x; i < n; i += stride) {
    const double sp = spot_at<kFp32>(paths, n, date, i);
    if (intrinsic(sp, k, is_put) > 0.0) {
      const double x = (sp - mean) / sd;
      const double y = cf[i];
      double xp = 1.0;
#pragma unroll
      for (int a = 0; a < kMaxMom; ++a) {
        if (a < nm) {
          acc[a] += xp;
          if (a < ncols) acc[kMaxMom + a] += y * xp;
          xp *= x;
        }
      }
    }
  }
  // pack: [0, nm) moments, [nm, nm+ncols) cross moments
  double packed[kMaxAccB];
#pragma unroll
  for (int a = 0; a < kMaxAccB; ++a) packed[a] = 0.0;
#pragma unroll
  for (int a = 0; a < kMaxMom; ++a)
    if (a < nm) packed[a] = acc[a];
#pragma unroll
  for (int a = 0; a < kMaxCols; ++a)
    if (a < ncols) packed[nm + a] = acc[kMaxMom + a];
  block_store<kMaxAccB>(packed, nm + ncols, part);
}

__global__ __launch_bounds__(kLsBlock) void lsmc_solve(const double* __restrict__ part,
                                                       const int n_blocks, const uint64_t date,
                                                       const int ncols,
                                                       DateState* __restrict__ ds) {
  DateState& d = ds[date];
  if (d.skip) return;
  const int nm = 2 * (ncols - 1) + 1;
  __shared__ double tot[kMaxAccB];
  reduce_block_partials(part, n_blocks, nm + ncols, tot);
  if (threadIdx.x == 0) {
    // Normal equations G c = b, G[i][j] = sum x^(i+j), b[i] = sum y x^i.
    double G[kMaxCols][kMaxCols], L[kMaxCols][kMaxCols], b[kMaxCols], w[kMaxCols];
    for (int i = 0; i < ncols; ++i) {
      for (int j = 0; j < ncols; ++j) G[i][j] = tot[i + j];
      b[i] = tot[nm + i];
    }
    // Cholesky G = L L^T (G is SPD when the ITM spots take >= ncols values)
    bool ok = true;
    for (int i = 0; i < ncols && ok; ++i) {
      for (int j = 0; j <= i; ++j) {
        double s = G[i][j];
        for (int q = 0; q < j; ++q) s -= L[i][q] * L[j][q];
        if (i == j) {
          if (!(s > 0.0)) {
            ok = false;
            break;
          }
          L[i][i] = sqrt(s);
        } else {
          L[i][j] = s / L[j][j];
        }
      }
    }
    if (!ok) {
      // Rank-deficient basis (all ITM spots equal): regress on the constant.
      for (int i = 0; i < ncols; ++i) d.coeff[i] = 0.0;
      d.coeff[0] = b[0] / G[0][0];
      return;
    }
    for (int i = 0; i < ncols; ++i) {
      double s = b[i];
      for (int q = 0; q < i; ++q) s -= L[i][q] * w[q];
      w[i] = s / L[i][i];
    }
    for (int i = ncols - 1; i >= 0; --i) {
      double s = w[i];
      for (int q = i + 1; q < ncols; ++q) s -= L[q][i] * d.coeff[q];
      d.coeff[i] = s / L[i][i];
    }
  }
}

// Final variance pass: sum (cf - mean)^2 (lsmc.rs:109).
__global__ __launch_bounds__(kLsBlock) void lsmc_var_pass(const double* __restrict__ cf,
                                                          const uint64_t n,
                                                          const double* __restrict__ mean_sum,
                                                          double* __restrict__ part) {
  const double mean = mean_sum[0] / (double)n;
  double acc[1] = {0.0};
  const uint64_t stride = (uint64_t)gridDim.x * kLsBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kLsBlock + threadIdx.x; i < n; i += stride) {
    const double dlt = cf[i] - mean;
    acc[0] += dlt * dlt;
  }
  block_store<1>(acc, 1, part);
}

__global__ void lsmc_reduce_scalar(const double* __restrict__ part, int n_blocks,
                                   double* __restrict__ out) {
  __shared__ double tot[1];
  reduce_block_partials(part, n_blocks, 1, tot);
  if (threadIdx.x == 0) out[0] = tot[0];
}

// ------------------------------------------------- persistent induction --
// ------------------------------------------------- one-launch induction --
// The whole backward induction in one cooperative launch of one 1024-thread
// block per CU.  Thread t owns paths t, t+G, t+2G, ... (G = grid threads, at
// most kPpt per thread): their cash flows and current spots stay in
// registers across dates, so a date costs one coalesced read of its spot
// column.  One grid barrier per date:
//   pass     exercise at date+1 (coefficients solved there), discount, and at
//            `date` the ITM count and raw moments of u = (S - c) / sc over ITM
//            paths (sum u^e, e < 2 deg + 1; sum cf u^e, e <= deg) -> block
//            partials; (c, sc) is the mean / std of the previous regression
//            date (K, K before the first), fixed before the pass
//   barrier
//   solve    every block sums the partials in the same fixed order, forms
//            mean and population std (floor 1e-12, lsmc.rs:67-71), maps the
//            moments to x = (S - mean) / std = al u + be by the binomial
//            expansion (al ~ 1, be ~ 0 since (c, sc) track the previous date)
//            and solves the normal equations (FP64 Cholesky) -- so every
//            block holds the same coefficients without a second barrier.
// The partials alternate between two buffers by date parity: a block may
// start writing the next date's partials while slower blocks still read this
// date's; the buffer it writes was last read before the barrier it passed.
constexpr int kCoopBlock = 1024;
constexpr int kCoopWaves = kCoopBlock / 64;
constexpr int kCoopMaxPpt = 16;

// Cross-block data of the cooperative kernel: each date's block partials.
// They go to slots of their own per date (never reused within a launch) that
// hold kEmptySlot when the kernel starts; a block stores its partial with a
// relaxed agent-scope store (coherent across the XCDs' L2s) and every block
// reads each slot with await_slot, which waits until the slot is filled.
// The slots are the flags: no grid barrier, no ordering between different
// locations, no L2 write-back or invalidate (cooperative groups' grid.sync()
// costs ~33 us per barrier at 256 blocks on MI355X).  Every block of the
// cooperative grid is resident, so every awaited slot gets filled.
template <class T>
__device__ __forceinline__ void st_agent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct LsmcCoopArgs {
  const void* paths;
  uint64_t n;
  uint32_t steps;
  int is_put, ncols;
  double k, discount;
  double* part;  // [steps + 1][kMaxAccB][grid] partial slots, empty at kernel start:
                 // slot `date` for dates 1..steps-1, 0 and `steps` for the two final sums
  double* out;   // [2]: sum cf, sum (cf - mean)^2
};

// Moments of x = (S - mean) / sdv = al u + be from the raw moments of
// u = (S - c) / sc (tot[q] = sum u^q, q < nm; tot[yoff + q] = sum cf u^q,
// q < ncols): sum x^e = sum_q C(e, q) al^q be^(e-q) tot[q], and the same for
// sum cf x^e with tot[yoff + q].  Lane e < nm of the calling wave writes g[e],
// lane 16 + e (e < ncols) writes h[e].  Per lane the sum is a Horner scheme
// in be over q = 0..e, one fused multiply-add per q on the statistics wave's
// serial path; C(e, q) = rint(e (e-1) ... (e-q+1) / q!) comes off a
// falling-factorial product (exact integers below 2^16; the rounded quotient
// lies within 1e-12 of the integer), which is not on that path.
__device__ __forceinline__ void moments_x(const double* tot, int ncols, int yoff, double al,
                                          double be, double* g, double* h) {
  constexpr double kInvFact[kMaxMom] = {1.0,       1.0,        1.0 / 2,     1.0 / 6,     1.0 / 24,
                                        1.0 / 120, 1.0 / 720,  1.0 / 5040,  1.0 / 40320};
  const int nm = 2 * (ncols - 1) + 1;
  int lane = threadIdx.x & 63;
  asm volatile("" : "+v"(lane));  // lane-dependent values are formed here, not kept live
  const bool is_g = lane < nm, is_h = lane >= 16 && lane < 16 + ncols;
  if (!is_g && !is_h) return;
  const int e = is_g ? lane : lane - 16;
  const double* src = is_g ? tot : tot + yoff;
  double s = 0.0, ff = 1.0, aq = 1.0;  // ff = e (e-1) ... (e-q+1), aq = al^q
#pragma unroll
  for (int q = 0; q < kMaxMom; ++q) {
    // every lane reads (index min(q, e) stays below nm / ncols): a read
    // under the q <= e condition would be a branch and an LDS wait per q
    const double v = src[q < e ? q : e];
    const double binom = rint(ff * kInvFact[q]);  // C(e, q)
    const double term = binom * aq * v;
    s = q <= e ? __builtin_fma(s, be, term) : s;
    ff *= (double)(e - q);
    aq *= al;
  }
  if (is_g)
    g[e] = s;
  else
    h[e] = s;
}

// A wave-uniform double moved to scalar registers.
__device__ __forceinline__ double uniform_d(double v) {
  const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
  const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}

// 1 / d from the hardware reciprocal and two Newton steps (within an ulp of
// the quotient; the solve's own rounding is coarser than that).
__device__ __forceinline__ double recip(double d) {
  double x = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, x, 1.0);
  x = __builtin_fma(x, e, x);
  e = __builtin_fma(-d, x, 1.0);
  return __builtin_fma(x, e, x);
}

// G c = h with G[i][j] = g[i + j] (i, j < ncols) in FP64: the least-squares
// minimiser of lsmc.rs:73-84 via the normal equations, solved by the calling
// wave (all 64 lanes).  Lane i holds row i of [G | h] (rows ncols..4 are
// identity rows).  Gauss-Jordan elimination without pivoting: step k
// broadcasts pivot row k with readlane and every other row subtracts its
// multiple of it, so afterwards lane i holds c_i = h_i / G_ii on its own and
// writes it.  G is SPD when the in-the-money spots take at least ncols
// values, so every pivot is positive; a pivot <= 0 marks a rank-deficient
// basis, and then the regression falls back to the constant.
__device__ __forceinline__ void solve_normal_eq_wave(const double* g, const double* h, int ncols,
                                                     double* coeff) {
  int lane = threadIdx.x & 63;
  asm volatile("" : "+v"(lane));  // lane-dependent values are formed here, not kept live
  const bool row = lane < ncols;
  double r[kMaxCols], hh = row ? h[lane] : 0.0;
#pragma unroll
  for (int j = 0; j < kMaxCols; ++j)
    r[j] = (row && j < ncols) ? g[(row ? lane : 0) + j] : (lane == j ? 1.0 : 0.0);
  bool ok = true;
  double my_inv = 1.0;  // lane i: 1 / (pivot i)
#pragma unroll
  for (int k = 0; k < kMaxCols; ++k) {
    // pivots past ncols are the identity rows': they leave rows < ncols alone
    if (k >= ncols) break;
    double pk[kMaxCols];
#pragma unroll
    for (int j = k; j < kMaxCols; ++j) pk[j] = readlane_d(r[j], k);
    const double phk = readlane_d(hh, k);
    ok = ok && pk[k] > 0.0;
    const double inv = recip(pk[k]);
    my_inv = lane == k ? inv : my_inv;
    const double f = lane != k ? r[k] * inv : 0.0;
#pragma unroll
    for (int j = k + 1; j < kMaxCols; ++j) r[j] -= f * pk[j];
    hh -= f * phk;
  }
  if (lane < kMaxCols) coeff[lane] = (ok && row) ? hh * my_inv : 0.0;
  if (!ok && lane == 0) coeff[0] = h[0] / g[0];  // rank-deficient basis: the constant
}

// ---- wave reduce-scatter of 16 FP64 slots (gfx950 lane exchanges) ----
// v[0..16) per lane -> lane 4k holds the wave total of slot k.  Each step
// halves the slots a lane carries and doubles the lanes summed into them:
//   permlane32_swap  lanes l, l^32      slots j | j+8  (by lane bit 5)
//   permlane16_swap  lanes l, l^16      slots j | j+4  (bit 4)
//   DPP row_mirror   lanes l, 15-l      slots j | j+2  (bit 3)
//   DPP half_mirror  lanes l, 7-l       slots 0 | 1    (bit 2)
//   DPP quad xor 2, xor 1               the same slot
// so lane l ends with slot 8 b5 + 4 b4 + 2 b3 + b2, summed over all 64 lanes
// in an order fixed by l (deterministic).
struct D2 {
  double a, b;
};
__device__ __forceinline__ D2 swap32(double a, double b) {
  const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(a), __double2loint(b), false, false);
  const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(a), __double2hiint(b), false, false);
  return D2{__hiloint2double((int)hi[0], (int)lo[0]), __hiloint2double((int)hi[1], (int)lo[1])};
}
__device__ __forceinline__ D2 swap16(double a, double b) {
  const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(a), __double2loint(b), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(a), __double2hiint(b), false, false);
  return D2{__hiloint2double((int)hi[0], (int)lo[0]), __hiloint2double((int)hi[1], (int)lo[1])};
}
template <int kCtrl>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_reduce_scatter16(const double (&v)[16]) {
  const unsigned l = threadIdx.x & 63u;
  double w[8], u[4], t[2];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const D2 r = swap32(v[j], v[j + 8]);
    w[j] = r.a + r.b;
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const D2 r = swap16(w[j], w[j + 4]);
    u[j] = r.a + r.b;
  }
  const bool b3 = (l >> 3) & 1u, b2 = (l >> 2) & 1u;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const double keep = b3 ? u[j + 2] : u[j], send = b3 ? u[j] : u[j + 2];
    t[j] = keep + dpp_f64<0x140>(send);  // row_mirror
  }
  const double keep = b2 ? t[1] : t[0], send = b2 ? t[0] : t[1];
  double r = keep + dpp_f64<0x141>(send);  // row_half_mirror
  r += dpp_f64<0x4E>(r);                   // quad_perm [2,3,0,1]
  r += dpp_f64<0xB1>(r);                   // quad_perm [1,0,3,2]
  return r;
}

// Block sums of acc[0..na) over the first kWaves waves (the other waves hold
// nothing): per-wave reduce-scatter, then the waves in order; thread a < na
// returns slot a's total (red is free again on return).  Called by every
// thread of the block.
template <int kWaves>
__device__ __forceinline__ double coop_block_sum(const double* acc, int na,
                                                 double (*red)[kMaxAccB]) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (wave < kWaves) {
    double v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = (e < kMaxAccB && e < na) ? acc[e < kMaxAccB ? e : 0] : 0.0;
    const double r = wave_reduce_scatter16(v);
    const int slot = lane >> 2;
    if ((lane & 3) == 0 && slot < na) red[wave][slot] = r;
  }
  lds_barrier();
  double s = 0.0;
  if (threadIdx.x < na)
    for (int w = 0; w < kWaves; ++w) s += red[w][threadIdx.x];
  lds_barrier();
  return s;
}

// The fixed-order total over blocks of column `col` (nb slots, one per
// block) for the calling wave: lane l adds blocks l, l + 64, ... (coalesced
// loads), then the lanes in a fixed butterfly.  The wave loads its slots
// together and re-loads all the still-empty ones together, so waiting for
// late blocks costs one load round trip per poll, not one per empty slot.
// The poll is bounded (2^24 rounds, ~0.5 s of sleeps): a slot that never
// fills yields a NaN total, never a hang.
__device__ __forceinline__ double column_total(const double* col, unsigned nb) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  for (unsigned b0 = 0; b0 < nb; b0 += 256) {  // 4 loads in flight per lane
    uint64_t bits[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {  // unpredicated loads (block 0 past nb)
      const unsigned b = b0 + 64 * k + lane;
      bits[k] = load_slot(&col[b < nb ? b : 0u]);
    }
    for (uint32_t round = 0;; ++round) {
      bool empty = false;
#pragma unroll
      for (int k = 0; k < 4; ++k) empty |= (b0 + 64 * k + lane < nb) & (bits[k] == kEmptySlot);
      if (!__builtin_amdgcn_ballot_w64(empty)) break;
      if (round == (1u << 24)) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
          bits[k] = bits[k] == kEmptySlot ? 0x7ff8000000000000ull : bits[k];  // NaN
        break;
      }
      __builtin_amdgcn_s_sleep(1);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const unsigned b = b0 + 64 * k + lane;
        if (bits[k] == kEmptySlot) bits[k] = load_slot(&col[b < nb ? b : 0u]);
      }
    }
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      v[k] = b0 + 64 * k + lane < nb ? __longlong_as_double((long long)bits[k]) : 0.0;
    s += (v[0] + v[1]) + (v[2] + v[3]);
  }
  return wave_sum_d(s);
}

// Every block: fixed-order totals over blocks of part[e * nb + b] (element e
// of block b's partial, e < na <= 16) -> tot (LDS); wave e sums element e.
// A block's na elements are stored by one store instruction, so once every
// block's element 0 is there the others are too, bar stragglers (which their
// waves' own polls wait for).  So wave 0 alone polls column 0 while the
// other waves wait at a block barrier, and then read their columns once:
// with all na waves polling together the polls' traffic slows the stores
// they wait for (measured on 256 blocks: 4.3 us per exchange, 2.4 us this way).
__device__ __forceinline__ void coop_total(const double* part, int na, double* tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned nb = gridDim.x;
  if (wave == 0) {
    const double s = column_total(part, nb);
    if (lane == 0) tot[0] = s;
  }
  lds_barrier();
  if (wave >= 1 && wave < na) {
    const double s = column_total(part + (uint64_t)wave * nb, nb);
    if (lane == 0) tot[wave] = s;
  }
  lds_barrier();
}

// Payoff max(sgn (s - k), 0) with sgn = -1 for a put, +1 for a call, and
// nsk = -sgn k: the fused multiply-add rounds once, so the value is k - s or
// s - k exactly as lsmc.rs:25-31 forms it.
__device__ __forceinline__ double intrinsic_sgn(double s, double sgn, double nsk) {
  const double v = __builtin_fma(sgn, s, nsk);
  return v > 0.0 ? v : 0.0;
}

// Exercise decision of lsmc.rs:91-98 with the coefficients of the last
// regression date (c[j] = 0 for j >= ncols); returns the (undiscounted) cash
// flow.  x is formed with the reciprocal of the std and the continuation
// value by Horner's rule in fused multiply-adds instead of sum c_j x^j: both
// move the value by a few ulps, below the difference between the normal
// equations here and the reference's QR, so no decision moves measurably.
__device__ __forceinline__ double exercise(double cf, double s, double sgn, double nsk,
                                           double mean, double inv_sdv, const double* c) {
  const double exv = intrinsic_sgn(s, sgn, nsk);
  if (exv > 0.0) {
    const double x = (s - mean) * inv_sdv;
    double cont = c[kMaxCols - 1];
#pragma unroll
    for (int j = kMaxCols - 2; j >= 0; --j) cont = __builtin_fma(cont, x, c[j]);
    if (exv > (cont > 0.0 ? cont : 0.0)) return exv;
  }
  return cf;
}

template <bool kFp32, int kPpt>
__global__ __launch_bounds__(kCoopBlock) void lsmc_coop_kernel(const LsmcCoopArgs a) {
  __shared__ double red[kCoopWaves][kMaxAccB];
  __shared__ double tot[kMaxAccB];
  __shared__ double coef_s[kMaxCols];
  __shared__ double g_s[kMaxMom], h_s[kMaxCols];
  __shared__ double ms_s[3];
  __shared__ int skip_s;
  // n <= kCoopMaxPpt x the grid's threads (< 2^32): path indices are 32-bit,
  // and a spot's address is a per-date column base plus a 32-bit offset.
  const uint32_t G = gridDim.x * kCoopBlock;
  const uint32_t t0 = blockIdx.x * kCoopBlock + threadIdx.x;
  // kPpt is the smallest power of two with kPpt x G >= n, so n > kPpt / 2 x G
  // and the first half of a thread's paths always exist: only the second
  // half needs the per-lane check (fewer live predicates).
  const uint64_t n_paths = a.n;
  auto live = [n_paths](int q, uint32_t i) { return q < kPpt / 2 || i < n_paths; };
  // Accumulators in a fixed layout: [0, kMaxMom) sum u^e and
  // [kMaxMom, kMaxAccB) sum cf u^e, all of them for every degree (static
  // register indices, no branches on the degree); the solve reads the first
  // 2 deg + 1 and deg + 1 of them.
  const int ncols = a.ncols, na = kMaxAccB;
  const bool deg4 = ncols == kMaxCols;
  const double sgn = a.is_put ? -1.0 : 1.0, nsk = -sgn * a.k;
  const uint64_t pstride = (uint64_t)gridDim.x * na;
  // Cash flows live in LDS (kPpt x 1024 doubles, <= 128 KB with one block
  // per CU).  The spots of date+1: FP32 paths above 4 per thread keep them
  // in registers as float (kRegSpot, with the next date's column loaded a
  // pass ahead); otherwise up to 8 per thread in LDS (kKeepSpot), and at 16
  // FP64 paths per thread they are read again (the previous date's column,
  // still in the caches).
  constexpr bool kRegSpot = kFp32 && kPpt > 4;
  constexpr bool kKeepSpot = !kRegSpot && kPpt <= 8;
  constexpr int kUnrollQ = (kPpt <= 8 || kRegSpot) ? kPpt : 4;  // per-path loops
  using SpotT = typename std::conditional<kFp32, float, double>::type;
  const uint32_t gstep = G * (uint32_t)sizeof(SpotT);  // bytes between a thread's paths
  __shared__ double cf[kPpt][kCoopBlock];
  __shared__ double sp[kKeepSpot ? kPpt : 1][kCoopBlock];
  SpotT cur[kRegSpot ? kPpt : 1];
  const int tx = threadIdx.x;
  const SpotCol<kFp32> last(a.paths, a.n, a.steps);
#pragma unroll kUnrollQ
  for (int q = 0; q < kPpt; ++q) {
    const uint32_t i = t0 + q * G;
    const SpotT r0 = last[i];  // 0 past n
    const double s0 = (double)r0;
    if constexpr (kKeepSpot) sp[q][tx] = s0;
    if constexpr (kRegSpot) cur[q] = r0;
    cf[q][tx] = live(q, i) ? intrinsic_sgn(s0, sgn, nsk) : 0.0;  // lsmc.rs:45-47
  }
  bool have_ex = false;  // a regression was solved at date+1
  double ex_mean = 0.0, ex_std = 1.0, ex_inv = 1.0, ex_c[kMaxCols];
  for (int j = 0; j < kMaxCols; ++j) ex_c[j] = 0.0;
  double c = a.k, sc = a.k;  // moment variable u = (S - c) / sc
  double inv_sc = 1.0 / a.k;
  // Up to 4 paths per thread, with kRegSpot, and at 16 FP64 paths per thread
  // (registers to spare there) the spots of the next date are loaded a pass
  // ahead, so their latency overlaps the solve and the exercise pass.
  constexpr bool kPrefetch = kPpt <= 4 || kRegSpot || kPpt == 16;
  SpotT nxt[kPrefetch ? kPpt : 1];
  if constexpr (kPrefetch) {
    const SpotCol<kFp32> first(a.paths, a.n, a.steps - 1);  // steps >= 2
#pragma unroll
    for (int q = 0; q < kPpt; ++q) {
      const uint32_t i = t0 + q * G;
      nxt[q] = first[i];  // 0 past n
    }
  }
  for (uint32_t date = a.steps - 1; date >= 1; --date) {
    double acc[kMaxAccB];
#pragma unroll
    for (int e = 0; e < kMaxAccB; ++e) acc[e] = 0.0;
    const SpotCol<kFp32> col0(a.paths, a.n, date), col1(a.paths, a.n, date + 1);
    const uint32_t toff = fresh_offset(t0, sizeof(SpotT));
    // exercise at date+1 with its regression, then one discount step
    // (lsmc.rs:91-98, 52-54); the spots of date are not needed yet, so
    // their prefetched loads are still landing
#pragma unroll kUnrollQ
    for (int q = 0; q < kPpt; ++q) {
      const uint32_t i = t0 + q * G;
      if (live(q, i)) {
        double v = cf[q][tx];
        if (have_ex) {
          const double s1 = kRegSpot    ? (double)cur[kRegSpot ? q : 0]
                            : kKeepSpot ? sp[kKeepSpot ? q : 0][tx]
                                        : (double)col1.at(toff + q * gstep);
          v = exercise(v, s1, sgn, nsk, ex_mean, ex_inv, ex_c);
        }
        cf[q][tx] = v * a.discount;
      }
    }
    // moments of the in-the-money spots of date (u basis)
#pragma unroll kUnrollQ
    for (int q = 0; q < kPpt; ++q) {
      const uint32_t i = t0 + q * G;
      if (live(q, i)) {
        const double y = cf[q][tx];
        const double s =
            kPrefetch ? (double)nxt[kPrefetch ? q : 0] : (double)col0.at(toff + q * gstep);
        if constexpr (kKeepSpot) sp[q][tx] = s;
        if constexpr (kRegSpot) cur[q] = nxt[q];
        if (intrinsic_sgn(s, sgn, nsk) > 0.0) {
          const double u = (s - c) * inv_sc;
          double up = 1.0, u4 = 0.0;
#pragma unroll
          for (int e = 0; e < kMaxMom - 2; ++e) {  // degree <= 3: u^0..u^6, cf u^0..u^3
            acc[e] += up;
            if (e < kMaxCols - 1) acc[kMaxMom + e] = __builtin_fma(y, up, acc[kMaxMom + e]);
            if (e == kMaxCols - 1) u4 = up;
            up *= u;
          }
          if (deg4) {  // block-uniform: u^7, u^8 and cf u^4
            acc[kMaxMom - 2] += up;
            acc[kMaxMom - 1] += up * u;
            acc[kMaxMom + kMaxCols - 1] = __builtin_fma(y, u4, acc[kMaxMom + kMaxCols - 1]);
          }
        }
      }
    }
    double* pp = a.part + (uint64_t)date * pstride;
    const double bs = coop_block_sum<kCoopWaves>(acc, na, red);
    // The storing wave waits for its store before it polls (its loads
    // would otherwise hold the store up: measured 2.9 -> 1.3 us per
    // one-word exchange); the other waves wait in coop_total anyway.
    if (threadIdx.x < na) st_agent(&pp[(uint64_t)threadIdx.x * gridDim.x + blockIdx.x], bs);
    if (threadIdx.x < 64) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    coop_total(pp, na, tot);
    // The spots of date-1 are loaded here: after the totals' loads (vector
    // memory returns in order), and ahead of the solve and the next exercise
    // pass, which wait on LDS only.
    if constexpr (kPrefetch) {
      if (date > 1) {
        const SpotCol<kFp32> colm(a.paths, a.n, date - 1);
        const uint32_t toffm = fresh_offset(t0, sizeof(SpotT));
#pragma unroll
        for (int q = 0; q < kPpt; ++q) nxt[q] = colm.at(toffm + q * gstep);  // 0 past n
      }
    }
    if (threadIdx.x < 64) {  // wave 0: statistics, the moments of x, the solve
      const double m = tot[0];
      const int skip = (uint64_t)m < (uint64_t)ncols + 1;  // lsmc.rs:62-64
      if (!skip) {
        const double inv_m = recip(m);
        const double mu = tot[1] * inv_m;
        double var = tot[2] * inv_m - mu * mu;
        var = var > 0.0 ? var : 0.0;
        const double mean = c + sc * mu;
        double sdv = sqrt(var) * sc;
        sdv = sdv > 1e-12 ? sdv : 1e-12;  // lsmc.rs:67-71
        const double inv_sdv = recip(sdv);
        moments_x(tot, ncols, kMaxMom, sc * inv_sdv, (c - mean) * inv_sdv, g_s, h_s);
        if (threadIdx.x == 0) {
          ms_s[0] = mean;
          ms_s[1] = sdv;
          ms_s[2] = inv_sdv;
        }
      }
      if (threadIdx.x == 0) skip_s = skip;
      if (!skip) {
        // g_s and h_s come from this wave's own lanes: one wave's LDS
        // operations complete in issue order, and the fence keeps the
        // compiler from moving the solve's reads above those writes.
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        solve_normal_eq_wave(g_s, h_s, ncols, coef_s);
      }
    }
    lds_barrier();
    have_ex = __builtin_amdgcn_readfirstlane(skip_s) == 0;  // the same in every block
    if (have_ex) {  // block-uniform values, held in scalar registers
      ex_mean = uniform_d(ms_s[0]);
      ex_std = uniform_d(ms_s[1]);
      ex_inv = uniform_d(ms_s[2]);
#pragma unroll
      for (int j = 0; j < kMaxCols; ++j) ex_c[j] = uniform_d(coef_s[j]);
      c = ex_mean;
      sc = ex_std;
      inv_sc = ex_inv;
    }
    if (date == 1) break;
  }
  // exercise at date 1, final discount (lsmc.rs:102-104), sum
  double acc1[kMaxAccB] = {};
  const SpotCol<kFp32> col1(a.paths, a.n, 1);
#pragma unroll kUnrollQ
  for (int q = 0; q < kPpt; ++q) {
    const uint32_t i = t0 + q * G;
    if (live(q, i)) {
      double v = cf[q][tx];
      if (have_ex) {
        const double s1 = kRegSpot    ? (double)cur[kRegSpot ? q : 0]
                          : kKeepSpot ? sp[kKeepSpot ? q : 0][tx]
                                      : (double)col1[i];
        v = exercise(v, s1, sgn, nsk, ex_mean, ex_inv, ex_c);
      }
      cf[q][tx] = v * a.discount;
      acc1[0] += cf[q][tx];
    }
  }
  double s1 = coop_block_sum<kCoopWaves>(acc1, 1, red);
  if (threadIdx.x == 0) st_agent(&a.part[blockIdx.x], s1);  // slot 0
  __syncthreads();
  coop_total(a.part, 1, tot);
  const double mean = tot[0] / (double)a.n;
  double acc2[kMaxAccB] = {};
#pragma unroll kUnrollQ
  for (int q = 0; q < kPpt; ++q) {
    const uint32_t i = t0 + q * G;
    if (live(q, i)) {
      const double dl = cf[q][tx] - mean;
      acc2[0] += dl * dl;
    }
  }
  double s2 = coop_block_sum<kCoopWaves>(acc2, 1, red);
  double* fin = a.part + (uint64_t)a.steps * pstride;  // slot `steps`
  if (threadIdx.x == 0) st_agent(&fin[blockIdx.x], s2);
  const double sum_cf = tot[0];
  __syncthreads();
  if (blockIdx.x == 0) {
    coop_total(fin, 1, tot);
    if (threadIdx.x == 0) {
      a.out[0] = sum_cf;
      a.out[1] = tot[0];
    }
  }
}

// Launches the cooperative induction with one block per CU when the paths
// fit kCoopMaxPpt per thread; *launched is false otherwise (the caller then
// runs the per-date kernels).
template <bool kFp32, int kPpt>
static int lsmc_coop_launch(const LsmcCoopArgs& args, unsigned grid, hipStream_t st,
                            const Err& err) {
  // the kernel's fit, asked once (same code object on every gfx950 device)
  static int bpc = -1;
  if (bpc < 0)
    STOCHA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, lsmc_coop_kernel<kFp32, kPpt>,
                                                            kCoopBlock, 0));
  if (bpc < 1) return err.set(STOCHA_EDEVICE, "LSMC kernel does not fit one block per CU");
  LsmcCoopArgs a = args;
  void* kargs[] = {&a};
  STOCHA_HIP(hipLaunchCooperativeKernel((const void*)lsmc_coop_kernel<kFp32, kPpt>, dim3(grid),
                                        dim3(kCoopBlock), kargs, 0, st));
  return STOCHA_OK;
}

template <bool kFp32>
static int lsmc_coop(const LsmcCoopArgs& args, unsigned grid, hipStream_t st, const Err& err) {
  const uint64_t G = (uint64_t)grid * kCoopBlock;
  const uint64_t ppt = (args.n + G - 1) / G;
  if (ppt <= 1) return lsmc_coop_launch<kFp32, 1>(args, grid, st, err);
  if (ppt <= 2) return lsmc_coop_launch<kFp32, 2>(args, grid, st, err);
  if (ppt <= 4) return lsmc_coop_launch<kFp32, 4>(args, grid, st, err);
  if (ppt <= 8) return lsmc_coop_launch<kFp32, 8>(args, grid, st, err);
  return lsmc_coop_launch<kFp32, 16>(args, grid, st, err);
}

template <bool kFp32>
static int lsmc_impl(const GbmConst& gc, const StreamDesc& sd, uint64_t n, uint64_t steps,
                     double k, double discount, int is_put, int ncols, double* price,
                     double* std_err, hipStream_t st, const Err& err) {
  if (steps > 0xffffffffull) return err.set(STOCHA_EINVAL, "steps must be below 2^32");
  const size_t esz = kFp32 ? sizeof(float) : sizeof(double);
  uint64_t nb64 = (n + kLsBlock - 1) / kLsBlock;
  const int nb = (int)(nb64 > 2048 ? 2048 : nb64);
  int dev = 0, cus = 0;
  STOCHA_HIP(hipGetDevice(&dev));
  STOCHA_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const bool coop = steps >= 2 && n <= (uint64_t)cus * kCoopBlock * kCoopMaxPpt &&
                    !getenv("STOCHA_LSMC_MULTIKERNEL");
  // workspace: paths | cf | partials | per-date state | 2 results (256-byte aligned);
  // the cooperative induction's partials are one slot set per date
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t part_n =
      coop ? (size_t)(steps + 1) * (size_t)cus * kMaxAccB : (size_t)nb * kMaxAccB;
  const size_t o_cf = up(esz * n * steps);
  const size_t o_part = o_cf + up(sizeof(double) * n);
  const size_t o_ds = o_part + up(sizeof(double) * part_n);
  const size_t o_scal = o_ds + up(sizeof(DateState) * (steps + 1));
  Workspace* ws = nullptr;
  if (int rc = workspace(o_scal + 256, 2 * sizeof(double), &ws, err)) return rc;
  char* base = static_cast<char*>(ws->dev);
  void* paths = base;
  double* cf = reinterpret_cast<double*>(base + o_cf);
  double* part = reinterpret_cast<double*>(base + o_part);
  DateState* dsp = reinterpret_cast<DateState*>(base + o_ds);
  double* scal = reinterpret_cast<double*>(base + o_scal);

  KernelTimer timer(st);
  timer.start();
  // the cooperative induction takes its cash flows from the last column
  // itself (no cf array) and starts from empty partial slots
  if (coop)
    if (int rc = fill_empty_slots(part, part_n, st, err)) return rc;
  {
    constexpr int kGb = gen_block<kFp32>();
    const uint64_t gb64 = (n + kGb - 1) / kGb;
    const int gnb = (int)(gb64 > 2048 ? 2048 : gb64);
    if (n * esz <= 0x7fffffffull)
      lsmc_gen_kernel<kFp32, true><<<gnb, kGb, 0, st>>>(gc, sd, n, (uint32_t)steps, paths,
                                                        coop ? nullptr : cf, k, is_put);
    else
      lsmc_gen_kernel<kFp32, false><<<gnb, kGb, 0, st>>>(gc, sd, n, (uint32_t)steps, paths,
                                                         coop ? nullptr : cf, k, is_put);
  }
  STOCHA_HIP(hipGetLastError());
  if (coop) {
    LsmcCoopArgs ca;
    ca.paths = paths;
    ca.n = n;
    ca.steps = (uint32_t)steps;
    ca.is_put = is_put;
    ca.ncols = ncols;
    ca.k = k;
    ca.discount = discount;
    ca.part = part;
    // the two totals go straight into the pinned host buffer (device-visible
    // through unified addressing; read after the stream sync)
    ca.out = ws->host;
    if (int rc = lsmc_coop<kFp32>(ca, (unsigned)cus, st, err)) return rc;
    STOCHA_HIP(hipGetLastError());
  } else {
    STOCHA_HIP(hipMemsetAsync(dsp, 0, sizeof(DateState) * (steps + 1), st));
    uint64_t ex_date = 0;
    for (uint64_t date = steps - 1; date >= 1; --date) {
      lsmc_pass_a<kFp32><<<nb, kLsBlock, 0, st>>>(paths, cf, n, ex_date, dsp, date, k, is_put,
                                                  ncols, discount, part);
      lsmc_reduce_a<<<1, kLsBlock, 0, st>>>(part, nb, date, k, ncols, dsp);
      lsmc_pass_b<kFp32><<<nb, kLsBlock, 0, st>>>(paths, cf, n, date, dsp, k, is_put, ncols,
                                                  part);
      lsmc_solve<<<1, kLsBlock, 0, st>>>(part, nb, date, ncols, dsp);
      ex_date = date;
      if (date == 1) break;
    }
    // exercise at the last regression date, final discount, sum of cf
    lsmc_pass_a<kFp32><<<nb, kLsBlock, 0, st>>>(paths, cf, n, ex_date, dsp, 0, k, is_put, ncols,
                                                discount, part);
    lsmc_reduce_scalar<<<1, kLsBlock, 0, st>>>(part, nb, scal);
    lsmc_var_pass<<<nb, kLsBlock, 0, st>>>(cf, n, scal, part);
    lsmc_reduce_scalar<<<1, kLsBlock, 0, st>>>(part, nb, scal + 1);
    STOCHA_HIP(hipGetLastError());
  }
  timer.stop(coop ? 2 : 0);
  if (!coop)
    STOCHA_HIP(hipMemcpyAsync(ws->host, scal, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
  const double mean = ws->host[0] / (double)n;
  *price = mean;
  *std_err = n <= 1 ? 0.0 : sqrt((ws->host[1] / (double)(n - 1)) / (double)n);
  return STOCHA_OK;
}

int run_lsmc(const GbmConst& gc, const StreamDesc& sd, uint64_t n_paths, uint64_t steps,
             double k, double discount, int is_put, int ncols, int precision, double* price,
             double* std_err, hipStream_t st, const Err& err) {
  if (precision == STOCHA_FP32)
    return lsmc_impl<true>(gc, sd, n_paths, steps, k, discount, is_put, ncols, price, std_err,
                           st, err);
  return lsmc_impl<false>(gc, sd, n_paths, steps, k, discount, is_put, ncols, price, std_err, st,
                          err);
}

}  // namespace stocha
