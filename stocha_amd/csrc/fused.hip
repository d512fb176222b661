// fused.hip -- terminal-value kernels that never store paths: European calls
// priced in-register at up to 8 strikes, and the pathwise delta/vega sums of
// greeks.rs:360-424.
//
// Each thread steps whole paths (grid-stride over path indices, every path on
// its own block-split stream), evaluates the payoff on S_T and accumulates
// FP64 sums.  Reduction: butterfly over the 64-lane wavefront, LDS across the
// block's waves, one partial per block; the last block to finish (a counter
// in the workspace) sums the partials in a fixed order -- thread t over
// blocks t, t+256, ..., then the same butterfly and wave order -- so the
// result does not depend on which block finishes last.  The grid size is a
// function of n_paths only, so the summation order -- and the result -- is
// the same on every run.
//
// Two schedules draw the same numbers:
//   kScheduleLoop  per-normal ziggurat loop (the reference's control flow)
//   kScheduleFeed  per-lane draw-feed state machine (zigsm.hpp) filling a
//                  small LDS queue of step inputs; the wave applies each step
//                  with all lanes together (terminal_feed_kernel).
#include <type_traits>
#include <utility>

#include "host.hpp"
#include "drivers.hpp"
#include "sums.hpp"

namespace stocha {

constexpr int kMaxStrikes = 8;
static_assert(kMaxAcc >= 2 * kMaxStrikes, "accumulators");
constexpr int kMaxScen = 8;  // bump scenarios per launch (greeks_fd)

// Payoffs write one path's contributions to the sums: v[0..n_acc()), zero for
// a lane without a path (valid false).
struct CallPayoff {
  double k[kMaxStrikes];
  int n;
  __device__ __forceinline__ int n_acc() const { return 2 * n; }
  __device__ __forceinline__ void terms(double st, bool valid, double* v) const {
#pragma unroll
    for (int i = 0; i < kMaxStrikes; ++i) {
      if (i < n) {
        double pay = st - k[i];
        pay = pay > 0.0 ? pay : 0.0;
        pay = valid ? pay : 0.0;
        v[2 * i] = pay;
        v[2 * i + 1] = pay * pay;
      }
    }
  }
};

// greeks.rs:388-418: per-path delta and vega terms for a call or put.
struct PathwisePayoff {
  double s0, strike, drift_t, sigma;  // drift_t = (r + 0.5*sigma*sigma)*t
  int is_call;
  __device__ __forceinline__ int n_acc() const { return 2; }
  __device__ __forceinline__ void terms(double s_t, bool valid, double* v) const {
    v[0] = 0.0;
    v[1] = 0.0;
    const bool in_money = valid && (is_call ? s_t > strike : s_t < strike);
    if (in_money) {
      const double sign = is_call ? 1.0 : -1.0;
      v[0] = sign * s_t / s0;
      const double log_return = log(s_t / s0);
      v[1] = sign * s_t * (log_return - drift_t) / sigma;
    }
  }
};

template <class Model, class Const, bool kFp32>
__device__ __forceinline__ double terminal_value(const Const& mc, const StreamDesc& sd,
                                                 uint64_t path, uint64_t steps,
                                                 const ZigLds* zig) {
  Pcg g{stream_at(&sd, path)};
  const U128 inc = sd.inc;
  Model m(mc, false);
#pragma unroll 1
  for (uint64_t j = 0; j < steps; ++j) (void)m.step(g, inc, zig);
  if constexpr (kFp32) {
    return mc.s0 * exp((double)m.ls);
  } else {
    return m.s;
  }
}

// Both schedules accumulate the same way: per round of paths (one path per
// lane), the payoff terms of the wave go into the wave's LDS row
// (wave_accumulate), so the loop and feed kernels sum in the same order and
// give the same bits.
template <class Model, class Const, bool kFp32, class Payoff>
__global__ __launch_bounds__(kFusedBlock) void terminal_payoff_kernel(
    const Const mc, const StreamDesc sd, const uint64_t n_paths, const uint64_t steps,
    const Payoff pay, double* __restrict__ partials, unsigned* __restrict__ counter,
    double* __restrict__ out) {
  __shared__ ZigLds zig;
  __shared__ double wacc[kFusedBlock / 64][kMaxAcc];
  zig_load(&zig);
  if (threadIdx.x < kFusedBlock / 64 * kMaxAcc) (&wacc[0][0])[threadIdx.x] = 0.0;
  __syncthreads();

  const uint64_t stride = (uint64_t)gridDim.x * kFusedBlock;
  for (uint64_t base = (uint64_t)blockIdx.x * kFusedBlock; base < n_paths; base += stride) {
    const uint64_t p = base + threadIdx.x;
    const bool valid = p < n_paths;
    double st = 0.0;
    if (valid) st = terminal_value<Model, Const, kFp32>(mc, sd, p, steps, &zig);
    double v[kMaxAcc];
    pay.terms(st, valid, v);
    wave_accumulate(v, pay.n_acc(), wacc);
  }
  finish_block_sums(wave_rows_total(pay.n_acc(), wacc), pay.n_acc(), wacc, partials, counter,
                    out);
}

// Queue kernel (the FEED schedule).  Each lane feeds its own draws into its
// model's source and appends the step inputs to a small ring in LDS (kQueue
// slots per lane, queue_path).  Before every step the wave draws until each
// lane holds one step's input -- lanes that already do keep drawing ahead
// while their ring has room -- and then all lanes apply the step together,
// so the step arithmetic never runs with half the wave masked.  Simulated on
// the draw-count distribution of the ziggurat, a ring of 8 Heston normals
// costs ~2.14 draws per step against the 2.06 minimum.
// Occupancy target per source (Src::kWaves): the FP32 GBM/Heston/Merton
// kernels run 7 waves per SIMD (<= 72 VGPRs, with two dwords spilled
// outside the path loop; 23,040 bytes of LDS per block, which seven blocks
// fit in 160 KB); the FP64 ones run 5 (87-91 VGPRs, no spills; 27,152 bytes
// of LDS, the 10 KB FP64 table block and a 16 KB ring).
template <class Src, class Model, class Const, class Payoff, int kQueue>
__global__ __launch_bounds__(kFusedBlock) __attribute__((amdgpu_waves_per_eu(
    Src::kWaves))) void terminal_feed_kernel(
    const Const mc, const StreamDesc sd, const uint64_t n_paths, const uint32_t steps,
    const Payoff pay, double* __restrict__ partials, unsigned* __restrict__ counter,
    double* __restrict__ out) {
  using Rec = typename Src::Rec;
  __shared__ typename Src::Tables zig;
  __shared__ Rec q[kQueue][kFusedBlock];
  __shared__ double wacc[kFusedBlock / 64][kMaxAcc];
  zig_load(&zig);
  if (threadIdx.x < kFusedBlock / 64 * kMaxAcc) (&wacc[0][0])[threadIdx.x] = 0.0;
  __syncthreads();

  const U128 inc = sd.inc;
  const uint64_t stride = (uint64_t)gridDim.x * kFusedBlock;
  for (uint64_t base = (uint64_t)blockIdx.x * kFusedBlock; base < n_paths; base += stride) {
    const uint64_t p = base + threadIdx.x;
    const bool valid = p < n_paths;
    Pcg g{stream_at(&sd, valid ? p : 0)};
    const double st = queue_path<Src, Model, kQueue, kFusedBlock>(mc, g, inc, steps, &zig, q);
    double v[kMaxAcc];
    pay.terms(st, valid, v);
    wave_accumulate(v, pay.n_acc(), wacc);
  }
  finish_block_sums(wave_rows_total(pay.n_acc(), wacc), pay.n_acc(), wacc, partials, counter,
                    out);
}

// Sums partials[b][i] over b: block i owns accumulator i; thread t adds
// blocks t, t+256, ... in order, then a fixed tree over the 256 threads, so
// the result does not depend on timing.
__global__ __launch_bounds__(256) void reduce_partials_kernel(const double* __restrict__ partials,
                                                              const int n_blocks, const int na,
                                                              double* __restrict__ out) {
  __shared__ double buf[256];
  const int i = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < n_blocks; b += 256) s += partials[(uint64_t)b * na + i];
  buf[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) buf[threadIdx.x] += buf[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[i] = buf[0];
}

int launch_reduce_partials(const double* part, int n_blocks, int na, double* out, hipStream_t st,
                           const Err& err) {
  reduce_partials_kernel<<<na, 256, 0, st>>>(part, n_blocks, na, out);
  STOCHA_HIP(hipGetLastError());
  return STOCHA_OK;
}

__global__ __launch_bounds__(256) void fill_empty_kernel(uint64_t* __restrict__ p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    p[i] = kEmptySlot;
}

int fill_empty_slots(double* p, size_t n, hipStream_t st, const Err& err) {
  size_t blocks = (n + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  if (blocks == 0) return STOCHA_OK;
  fill_empty_kernel<<<(unsigned)blocks, 256, 0, st>>>(reinterpret_cast<uint64_t*>(p), n);
  STOCHA_HIP(hipGetLastError());
  return STOCHA_OK;
}


template <class Model, class Src, int kQueue, class Const, bool kFp32, class Payoff>
static int run_terminal(const Const& mc, const StreamDesc& sd, uint64_t n_paths, uint64_t steps,
                        const Payoff& pay, int na, double* h_out, int schedule, hipStream_t st,
                        const Err& err, const ScratchTarget* into = nullptr) {
  if (steps > 0xffffffffull) return err.set(STOCHA_EINVAL, "steps must be below 2^32");
  const unsigned grid = fused_grid(n_paths);
  double* part;
  unsigned* counter;
  double* out;
  Workspace* ws = nullptr;
  if (into) {  // caller-placed partials, counter and totals (nothing allocated)
    if ((size_t)grid * na > into->max_partials)
      return err.set(STOCHA_EINVAL, "%u blocks x %d sums exceed the %zu partials of the scratch",
                     grid, na, into->max_partials);
    part = into->partials;
    counter = into->counter;
    out = into->totals;
  } else {
    if (int rc = workspace(0, sizeof(double) * kMaxAcc, &ws, err)) return rc;
    part = ws->sums_part;
    counter = ws->counter;
    // the last block writes the totals straight into the pinned host buffer
    // (device-visible through unified addressing; read after the stream sync)
    out = ws->host;
  }
  KernelTimer timer(st);
  timer.start();
  if (schedule == kScheduleLoop)
    terminal_payoff_kernel<Model, Const, kFp32, Payoff><<<grid, kFusedBlock, 0, st>>>(
        mc, sd, n_paths, steps, pay, part, counter, out);
  else
    terminal_feed_kernel<Src, Model, Const, Payoff, kQueue><<<grid, kFusedBlock, 0, st>>>(
        mc, sd, n_paths, (uint32_t)steps, pay, part, counter, out);
  STOCHA_HIP(hipGetLastError());
  timer.stop(1);
  if (into) return STOCHA_OK;
  STOCHA_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < na; ++i) h_out[i] = ws->host[i];
  return STOCHA_OK;
}

int launch_fused(int model, int precision, int schedule, const void* mconst,
                 const StreamDesc& sd, uint64_t n_paths, uint64_t steps, const double* strikes,
                 int n_strikes, double* h_sums, hipStream_t st, const Err& err,
                 const ScratchTarget* into) {
  CallPayoff pay{};
  pay.n = n_strikes;
  for (int i = 0; i < n_strikes; ++i) pay.k[i] = strikes[i];
  const int na = 2 * n_strikes;
  const bool f32 = precision == STOCHA_FP32;
  switch (model) {
    case kGbm: {
      const GbmConst& c = *(const GbmConst*)mconst;
      return f32 ? run_terminal<Gbm32, GbmSrc<true>, 16, GbmConst, true>(c, sd, n_paths, steps, pay, na, h_sums, schedule, st, err, into)
                 : run_terminal<Gbm64, GbmSrc<false>, 8, GbmConst, false>(c, sd, n_paths, steps, pay, na, h_sums, schedule, st, err, into);
    }
    case kHeston: {
      const HestonConst& c = *(const HestonConst*)mconst;
      return f32 ? run_terminal<Heston32, HestonSrc<true>, 16, HestonConst, true>(c, sd, n_paths, steps, pay, na, h_sums, schedule, st, err, into)
                 : run_terminal<Heston64, HestonSrc<false>, 8, HestonConst, false>(c, sd, n_paths, steps, pay, na, h_sums, schedule, st, err, into);
    }
    case kMerton: {
      const MertonConst& c = *(const MertonConst*)mconst;
      return f32 ? run_terminal<Merton32, MertonSrc<true, Merton32>, 8, MertonConst, true>(c, sd, n_paths, steps, pay, na, h_sums, schedule, st, err, into)
                 : run_terminal<Merton64, MertonSrc<false, Merton64>, 4, MertonConst, false>(c, sd, n_paths, steps, pay, na, h_sums, schedule, st, err, into);
    }
    default:
      return err.set(STOCHA_EINVAL, "unknown model id %d", model);
  }
}

// ---------------------------------------------------------------------------
// The no-allocation form of the fused European call
// (stocha_european_fused_noalloc).  Its only device memory beyond the
// caller's inputs and outputs is `scratch`, a file-scope device array holding
// the block-partial slots (fused_grid's at most 4096 blocks x kMaxAcc sums,
// set empty once per device at set-up and left empty by every launch), the
// totals and the last-block counter: 512.25 KiB on each device the library
// runs on, held until the process ends.  Host code reaches it only through
// the address hipGetSymbolAddress returns for the current device (kept per
// device); the array's name is a host-side stand-in, not a device address.
// Every call on a device shares that device's array, so calls that overlap on
// one device race.
constexpr size_t kScratchPartials = (size_t)4096 * kMaxAcc;
constexpr size_t kScratchTotals = kScratchPartials;           // kMaxAcc doubles
constexpr size_t kScratchCounter = kScratchTotals + kMaxAcc;  // one uint32 (a double slot)
constexpr size_t kScratchDoubles = kScratchCounter + 16;
static __device__ __align__(256) double scratch[kScratchDoubles];
static_assert(kScratchDoubles * sizeof(double) == 524544, "stated scratch size");

constexpr int kScratchDevices = 64;

// Device address of `scratch` on the current device, looked up once per
// device (callers hold the library mutex).
static int scratch_address(double** out, const Err& err) {
  static double* addr[kScratchDevices] = {};
  int dev = 0;
  STOCHA_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices)
    return err.set(STOCHA_EDEVICE, "device %d is beyond the %d devices of the scratch table", dev,
                   kScratchDevices);
  if (!addr[dev]) {
    void* p = nullptr;
    STOCHA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return err.set(STOCHA_EDEVICE, "no device address for the scratch array");
    // the partial slots start empty; every reduction leaves them so
    if (int rc = fill_empty_slots(static_cast<double*>(p), kScratchPartials, nullptr, err))
      return rc;
    STOCHA_HIP(hipStreamSynchronize(nullptr));
    addr[dev] = static_cast<double*>(p);
  }
  *out = addr[dev];
  return STOCHA_OK;
}

__global__ void clear_counter_kernel(unsigned* __restrict__ counter) {
  if (threadIdx.x == 0) *counter = 0u;
}

// Prices, standard errors and sums from the totals, with the host's
// expressions of stocha_european_fused (same operations, same order).
__global__ void fused_epilogue_kernel(const double* __restrict__ tot, const int n_strikes,
                                      const double disc, const double n,
                                      double* __restrict__ price, double* __restrict__ std_err,
                                      double* __restrict__ sums) {
  const int k = threadIdx.x;
  if (k >= n_strikes) return;
  const double s1 = tot[2 * k], s2 = tot[2 * k + 1];
  price[k] = disc * s1 / n;
  double var = n > 1 ? (s2 - s1 * s1 / n) / (n - 1.0) : 0.0;
  var = var > 0.0 ? var : 0.0;
  std_err[k] = disc * sqrt(var / n);
  if (sums) {
    sums[2 * k] = s1;
    sums[2 * k + 1] = s2;
  }
}

int launch_fused_noalloc(int model, int precision, int schedule, const void* mconst,
                         const StreamDesc& sd, uint64_t n_paths, uint64_t steps,
                         const double* strikes, int n_strikes, double disc, double* d_price,
                         double* d_std_err, double* d_sums, hipStream_t st, const Err& err) {
  double* base = nullptr;
  if (int rc = scratch_address(&base, err)) return rc;
  ScratchTarget into;
  into.partials = base;
  into.max_partials = kScratchPartials;
  into.totals = base + kScratchTotals;
  into.counter = reinterpret_cast<unsigned*>(base + kScratchCounter);
  // the counter must start at zero (a fill kernel rather than a memset node)
  clear_counter_kernel<<<1, 64, 0, st>>>(into.counter);
  STOCHA_HIP(hipGetLastError());
  if (int rc = launch_fused(model, precision, schedule, mconst, sd, n_paths, steps, strikes,
                            n_strikes, nullptr, st, err, &into))
    return rc;
  fused_epilogue_kernel<<<1, 64, 0, st>>>(into.totals, n_strikes, disc, (double)n_paths, d_price,
                                          d_std_err, d_sums);
  STOCHA_HIP(hipGetLastError());
  STOCHA_HIP(hipStreamSynchronize(st));
  return STOCHA_OK;
}

int run_greeks_pathwise(const GbmConst& gc, const StreamDesc& sd, uint64_t n_paths,
                        uint64_t steps, double s0, double r, double sigma, double t,
                        double strike, int is_call, double* delta_sum, double* vega_sum,
                        hipStream_t st, const Err& err) {
  PathwisePayoff pay{};
  pay.s0 = s0;
  pay.strike = strike;
  pay.drift_t = (r + 0.5 * sigma * sigma) * t;  // greeks.rs:411
  pay.sigma = sigma;
  pay.is_call = is_call;
  double h[2];
  int rc = run_terminal<Gbm64, GbmSrc<false>, 8, GbmConst, false>(gc, sd, n_paths, steps, pay, 2, h,
                                                               kScheduleFeed, st, err);
  *delta_sum = h[0];
  *vega_sum = h[1];
  return rc;
}

// ---------------------------------------------------------------------------
// Bump-and-revalue Greeks (greeks.rs:227-358) on shared draws.  Every
// scenario of greeks_fd_core prices the same paths with the same seed and
// block size (common random numbers), and a bump of s0, sigma / v0, mu or t
// leaves the sequence of draws a path consumes unchanged -- except Merton's t
// bump, which moves lambda*dt and so which uniforms jump, and, under the
// Heston QE scheme, the v0 and t bumps, which move the variance path and so
// whether a step's first draw is a normal or a uniform.  One launch
// therefore steps up to kMaxScen scenarios of a path on one stream of draws:
// the queue entries are produced once and applied to every scenario's model,
// each with its own constants, in the reference's FP64 arithmetic, so every
// scenario's S_T is the value its own path function would give.
template <class Const>
struct FdScenarios {
  Const c[kMaxScen];
  int n;
  double strike;
  double sign;  // payoff max(sign * (S_T - strike), 0): call +1, put -1 (greeks.rs:12-17)
};

template <class Model, int kMax>
struct ModelSet {
  Model m[kMax];
  int n;
  template <class Const, size_t... I>
  __device__ __forceinline__ ModelSet(const Const* c, int n_, std::index_sequence<I...>)
      : m{Model(c[I], false)...}, n(n_) {}
  template <class... A>
  __device__ __forceinline__ void apply(A... a) {
#pragma unroll
    for (int s = 0; s < kMax; ++s)
      if (s < n) m[s].apply(a...);
  }
  template <class... A>
  __device__ __forceinline__ void apply_jr(A... a) {
#pragma unroll
    for (int s = 0; s < kMax; ++s)
      if (s < n) m[s].apply_jr(a...);
  }
  template <class... A>
  __device__ __forceinline__ void apply_qe(A... a) {
#pragma unroll
    for (int s = 0; s < kMax; ++s)
      if (s < n) m[s].apply_qe(a...);
  }
};

// Ring slots per lane for the FP64 sources: 8 normals, or 4 Merton entries
// (16 bytes each), 16 KB per block either way.
template <class Src>
constexpr int kRing64 = sizeof(typename Src::Rec) > 8 ? 4 : 8;

// kWaves: 3 for Heston (two FP64 state words per scenario), else 4
template <class Src, class Model, class Const, int kQueue, int kWaves>
__global__ __launch_bounds__(kFusedBlock) __attribute__((amdgpu_waves_per_eu(kWaves))) void
fd_feed_kernel(const FdScenarios<Const> sc, const StreamDesc sd, const uint64_t n_paths,
               const uint32_t steps, double* __restrict__ partials,
               unsigned* __restrict__ counter, double* __restrict__ out) {
  using Rec = typename Src::Rec;
  __shared__ typename Src::Tables zig;
  __shared__ Rec q[kQueue][kFusedBlock];
  __shared__ double red[kFusedBlock / 64][kMaxAcc];
  zig_load(&zig);
  __syncthreads();

  double acc[kMaxAcc];
#pragma unroll
  for (int i = 0; i < kMaxAcc; ++i) acc[i] = 0.0;

  const U128 inc = sd.inc;
  const uint64_t stride = (uint64_t)gridDim.x * kFusedBlock;
  for (uint64_t base = (uint64_t)blockIdx.x * kFusedBlock; base < n_paths; base += stride) {
    const uint64_t p = base + threadIdx.x;
    const bool valid = p < n_paths;
    Pcg g{stream_at(&sd, valid ? p : 0)};
    Src src(sc.c[0]);  // the draw grammar is the same in every scenario of the launch
    ModelSet<Model, kMaxScen> ms(sc.c, sc.n, std::make_index_sequence<kMaxScen>());
    queue_run<Src, kQueue, kFusedBlock>(src, ms, g, inc, steps, &zig, q);
    if (valid) {
#pragma unroll
      for (int s = 0; s < kMaxScen; ++s) {
        if (s < sc.n) {
          double pay = sc.sign * (ms.m[s].terminal() - sc.strike);
          acc[s] += pay > 0.0 ? pay : 0.0;
        }
      }
    }
  }
  finish_sums(acc, sc.n, red, partials, counter, out);
}

template <class Model, class Src, class Const>
static int run_fd(const Const* c, int n, double strike, double sign, const StreamDesc& sd,
                  uint64_t n_paths, uint64_t steps, double* h_sums, hipStream_t st,
                  const Err& err) {
  if (steps > 0xffffffffull) return err.set(STOCHA_EINVAL, "steps must be below 2^32");
  if (n < 1 || n > kMaxScen) return err.set(STOCHA_EINVAL, "1 to %d scenarios per launch", kMaxScen);
  FdScenarios<Const> sc{};
  for (int i = 0; i < n; ++i) sc.c[i] = c[i];
  for (int i = n; i < kMaxScen; ++i) sc.c[i] = c[0];
  sc.n = n;
  sc.strike = strike;
  sc.sign = sign;
  const unsigned grid = fused_grid(n_paths);
  Workspace* ws = nullptr;
  if (int rc = workspace(0, sizeof(double) * kMaxAcc, &ws, err)) return rc;
  KernelTimer timer(st);
  timer.start();
  constexpr int kWaves =
      std::is_same<Model, Heston64>::value || std::is_same<Model, HestonQe64>::value ? 3 : 4;
  fd_feed_kernel<Src, Model, Const, kRing64<Src>, kWaves><<<grid, kFusedBlock, 0, st>>>(
      sc, sd, n_paths, (uint32_t)steps, ws->sums_part, ws->counter, ws->host);
  STOCHA_HIP(hipGetLastError());
  timer.stop(1);
  STOCHA_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < n; ++i) h_sums[i] = ws->host[i];
  return STOCHA_OK;
}

int launch_fd(int model, const void* consts, int n, double strike, double sign,
              const StreamDesc& sd, uint64_t n_paths, uint64_t steps, double* h_sums,
              hipStream_t st, const Err& err) {
  switch (model) {
    case kGbm:
      return run_fd<Gbm64, GbmSrc<false>>(static_cast<const GbmConst*>(consts), n, strike, sign,
                                          sd, n_paths, steps, h_sums, st, err);
    case kHeston:
      return run_fd<Heston64, HestonSrc<false>>(static_cast<const HestonConst*>(consts), n,
                                                strike, sign, sd, n_paths, steps, h_sums, st,
                                                err);
    case kMerton:
      return run_fd<Merton64, MertonSrc<false, Merton64>>(static_cast<const MertonConst*>(consts),
                                                          n, strike, sign, sd, n_paths, steps,
                                                          h_sums, st, err);
    case kHestonQe:
      return run_fd<HestonQe64, HestonQeSrc<HestonQe64>>(
          static_cast<const HestonQeConst*>(consts), n, strike, sign, sd, n_paths, steps, h_sums,
          st, err);
    default:
      return err.set(STOCHA_EINVAL, "unknown model id %d", model);
  }
}

// greeks.rs:91-167 (ModelSpec::terminal_prices): S_T of every path, FP64, in
// the reference's arithmetic (the custom-payoff form of greeks_fd).
template <class Src, class Model, class Const>
__global__ __launch_bounds__(kFusedBlock) __attribute__((amdgpu_waves_per_eu(4))) void
terminals_kernel(const Const mc, const StreamDesc sd, const uint64_t n_paths,
                 const uint32_t steps, double* __restrict__ out) {
  using Rec = typename Src::Rec;
  constexpr int kQueue = kRing64<Src>;
  __shared__ typename Src::Tables zig;
  __shared__ Rec q[kQueue][kFusedBlock];
  zig_load(&zig);
  __syncthreads();
  const uint64_t stride = (uint64_t)gridDim.x * kFusedBlock;
  for (uint64_t base = (uint64_t)blockIdx.x * kFusedBlock; base < n_paths; base += stride) {
    const uint64_t p = base + threadIdx.x;
    const bool valid = p < n_paths;
    Pcg g{stream_at(&sd, valid ? p : 0)};
    const double st = queue_path<Src, Model, kQueue, kFusedBlock>(mc, g, sd.inc, steps, &zig, q);
    if (valid) out[p] = st;
  }
}

int launch_terminals(int model, const void* mconst, const StreamDesc& sd, uint64_t n_paths,
                     uint64_t steps, double* d_out, hipStream_t st, const Err& err) {
  if (steps > 0xffffffffull) return err.set(STOCHA_EINVAL, "steps must be below 2^32");
  if (n_paths == 0) return STOCHA_OK;
  const unsigned grid = fused_grid(n_paths);
  KernelTimer timer(st);
  timer.start();
  switch (model) {
    case kGbm:
      terminals_kernel<GbmSrc<false>, Gbm64><<<grid, kFusedBlock, 0, st>>>(
          *static_cast<const GbmConst*>(mconst), sd, n_paths, (uint32_t)steps, d_out);
      break;
    case kHeston:
      terminals_kernel<HestonSrc<false>, Heston64><<<grid, kFusedBlock, 0, st>>>(
          *static_cast<const HestonConst*>(mconst), sd, n_paths, (uint32_t)steps, d_out);
      break;
    case kMerton:
      terminals_kernel<MertonSrc<false, Merton64>, Merton64><<<grid, kFusedBlock, 0, st>>>(
          *static_cast<const MertonConst*>(mconst), sd, n_paths, (uint32_t)steps, d_out);
      break;
    case kHestonQe:
      terminals_kernel<HestonQeSrc<HestonQe64>, HestonQe64><<<grid, kFusedBlock, 0, st>>>(
          *static_cast<const HestonQeConst*>(mconst), sd, n_paths, (uint32_t)steps, d_out);
      break;
    default:
      return err.set(STOCHA_EINVAL, "unknown model id %d", model);
  }
  STOCHA_HIP(hipGetLastError());
  timer.stop(1);
  return STOCHA_OK;
}

}  // namespace stocha
