// host.hpp -- host-side plumbing shared by the entry points: error strings,
// stream-table setup (128-bit jump-ahead on the host), device buffers and the
// kernel-time probe.
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include <mutex>
#include <vector>

#include "models.hpp"
#include "stocha_gpu.h"

namespace stocha {

// ---------------------------------------------------------------- errors --
struct Err {
  char* buf;
  size_t len;
  int set(int code, const char* fmt, ...) const __attribute__((format(printf, 3, 4))) {
    if (buf && len) {
      va_list ap;
      va_start(ap, fmt);
      vsnprintf(buf, len, fmt, ap);
      va_end(ap);
    }
    return code;
  }
};

#define STOCHA_HIP(call)                                                                  \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return err.set(e_ == hipErrorOutOfMemory ? STOCHA_ENOMEM : STOCHA_EDEVICE,          \
                     "HIP error %s at %s:%d (%s)", hipGetErrorName(e_), __FILE__, __LINE__, \
                     #call);                                                              \
  } while (0)

// ------------------------------------------------------- global state --
struct Globals {
  std::mutex mu;
  hipStream_t stream = nullptr;
  bool timing = false;
  double last_ms = 0.0;
  uint64_t last_launches = 0;
  bool timer_pending = false;  // ev0..ev1 recorded, last_ms not yet read from them
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};
Globals& globals();

// ------------------------------------------------------ device buffers --
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t n, const Err& err) {
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
      p = nullptr;
      return err.set(STOCHA_ENOMEM, "device allocation of %zu bytes failed (%s)", n,
                     hipGetErrorName(e));
    }
    bytes = n;
    return STOCHA_OK;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// ----------------------------------------------------- workspaces --
// Per-device buffers reused across calls instead of a hipMalloc/hipFree pair
// per call (hipFree waits for the whole device).  Calls are serialized by
// Globals::mu and have finished on the device when they return, so one set
// per device suffices.  Buffers grow on demand and live until process exit.
//   dev     device scratch (block partials, results, paths)
//   counter kCounters zeroed uint32s: [0] the last-block reductions of the
//           fused kernels; every kernel leaves them as it found them
//   sums_part  the fused reductions' block-partial slots (sums.hpp), empty
//           between launches
//   host    pinned host memory for results copied back
// A loud error when no device is visible (the library has no CPU path).
inline int check_device(const Err& err) {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return err.set(STOCHA_EDEVICE, "no HIP device available (%s)",
                   e == hipSuccess ? "0 devices" : hipGetErrorName(e));
  return STOCHA_OK;
}

// ------------------------------------------------------- output checks --
// n * m * elem, false when the product does not fit in 64 bits.
inline bool byte_count(uint64_t n, uint64_t m, uint64_t elem, size_t* bytes) {
  uint64_t nm = 0, b = 0;
  if (__builtin_mul_overflow(n, m, &nm) || __builtin_mul_overflow(nm, elem, &b)) return false;
  *bytes = (size_t)b;
  return true;
}

// `bytes` bytes from p on must lie in one device allocation of the current
// device: the kernels write device outputs directly, and a host pointer, a
// pointer on another device or a short buffer would fault the GPU or write
// past the caller's memory.
inline int check_device_span(const void* p, size_t bytes, const char* name, const Err& err) {
  if (!p) return err.set(STOCHA_EINVAL, "%s must not be NULL", name);
  int dev = 0;
  STOCHA_HIP(hipGetDevice(&dev));
  hipPointerAttribute_t a{};
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // keep the failed query out of the next launch check
    return err.set(STOCHA_EINVAL, "%s is not device memory of the current device", name);
  }
  if (a.type != hipMemoryTypeDevice || a.device != dev)
    return err.set(STOCHA_EINVAL,
                   "%s is not device memory of the current device (memory type %d, device %d; "
                   "current device %d)", name, (int)a.type, a.device, dev);
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  e = hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return err.set(STOCHA_EINVAL, "%s: the allocation it points into is unknown", name);
  }
  const uintptr_t lo = (uintptr_t)base, q = (uintptr_t)p;
  if (q < lo || q - lo > size || bytes > size - (q - lo))
    return err.set(STOCHA_EINVAL, "%s holds %zu bytes from the pointer on, the result needs %zu",
                   name, q < lo || q - lo > size ? (size_t)0 : size - (size_t)(q - lo), bytes);
  return STOCHA_OK;
}

// An output of n x m values of elem bytes placed by `where`: STOCHA_HOST or
// STOCHA_DEVICE only, a byte count that fits, and for STOCHA_DEVICE a device
// buffer of the current device that holds it.  *bytes = the byte count.
inline int check_output(const void* out, int where, uint64_t n, uint64_t m, uint64_t elem,
                        const char* name, size_t* bytes, const Err& err) {
  if (where != STOCHA_HOST && where != STOCHA_DEVICE)
    return err.set(STOCHA_EINVAL, "where must be STOCHA_HOST (0) or STOCHA_DEVICE (1), got %d",
                   where);
  if (!byte_count(n, m, elem, bytes))
    return err.set(STOCHA_EINVAL, "%s of %llu x %llu values does not fit in memory", name,
                   (unsigned long long)n, (unsigned long long)m);
  if (where == STOCHA_DEVICE && *bytes) {
    if (int rc = check_device(err)) return rc;
    return check_device_span(out, *bytes, name, err);
  }
  return STOCHA_OK;
}

struct Workspace {
  void* dev = nullptr;
  size_t dev_bytes = 0;
  unsigned* counter = nullptr;
  double* sums_part = nullptr;  // kSumsPartials slots, empty between launches
  double* host = nullptr;
  size_t host_bytes = 0;
};
// n slots set to the empty-slot pattern of the fused reductions (a fill
// kernel on st).
int fill_empty_slots(double* p, size_t n, hipStream_t st, const Err& err);
constexpr int kCounters = 1024;
// Workspace of the current device with at least the given sizes.
int workspace(size_t dev_bytes, size_t host_bytes, Workspace** out, const Err& err);

// ------------------------------------------------- kernel-time probe --
// Records events around the dominant kernel(s) of a call on the work stream.
struct KernelTimer {
  bool on;
  hipStream_t s;
  explicit KernelTimer(hipStream_t st) : on(globals().timing), s(st) {}
  void start() {
    if (!on) return;
    globals().timer_pending = false;
    globals().last_ms = 0.0;
    (void)hipEventRecord(globals().ev0, s);
  }
  // The elapsed time is read when it is asked for (stocha_last_kernel_ms),
  // by then normally complete: the entry point's own stream sync is the only
  // wait inside the call.
  void stop(uint64_t launches) {
    if (!on) return;
    (void)hipEventRecord(globals().ev1, s);
    globals().timer_pending = true;
    globals().last_launches = launches;
  }
};

// ------------------------------------------------------- stream tables --
// Affine map of the LCG after `steps` steps: s -> a*s + c (mod 2^128).
typedef unsigned __int128 u128h;
inline void lcg_power(u128h inc, u128h steps, u128h* a_out, u128h* c_out) {
  u128h acc_a = 1, acc_c = 0, cur_a = kPcgMul, cur_c = inc;
  while (steps) {
    if (steps & 1) {
      acc_a = acc_a * cur_a;
      acc_c = acc_c * cur_a + cur_c;
    }
    cur_c = (cur_a + 1) * cur_c;
    cur_a = cur_a * cur_a;
    steps >>= 1;
  }
  *a_out = acc_a;
  *c_out = acc_c;
}

inline U128 to_u128(u128h v) { return U128{(uint64_t)v, (uint64_t)(v >> 64)}; }

// Pcg64Dxsm::new(seed) (pcg64.rs:23-30) and jumps of block*2^k steps.
inline void make_stream_desc(uint64_t seed, uint64_t block, StreamDesc* d) {
  u128h s = seed;
  u128h stream = s + (((u128h)0xda3e39cb94b95bdbull << 64) | 0x7ef55da5c0374f5aull);
  u128h inc = (stream << 1) | 1;
  u128h state = s + inc;
  state = state * (u128h)kPcgMul + inc;
  d->base = to_u128(state);
  d->inc = to_u128(inc);
  u128h a, c;
  lcg_power(inc, (u128h)block, &a, &c);
  for (int k = 0; k < kJumpBits; ++k) {
    d->jump[k].a = to_u128(a);
    d->jump[k].c = to_u128(c);
    // square the map: (a, c) o (a, c) = (a^2, a*c + c)
    c = a * c + c;
    a = a * a;
  }
}

// ------------------------------------------------------ constant setup --
// The expressions below are the reference's, evaluated on the host in
// binary64 (this file is compiled with -ffp-contract=off).
GbmConst gbm_const(double s0, double mu, double sigma, double t, uint64_t steps, double q);
HestonConst heston_const(const double* p7, double t, uint64_t steps);
HestonQeConst heston_qe_const(const double* p7, double t, uint64_t steps);
MertonConst merton_const(const double* p6, double t, uint64_t steps);

// ------------------------------------------------------------ launchers --
// paths.hip
int launch_paths(int model, const void* mconst, const StreamDesc& sd, uint64_t n_rows,
                 uint64_t base_rows, uint64_t steps, double* d_out, hipStream_t st,
                 const Err& err);
// fused.hip
enum Schedule : int { kScheduleFeed = 0, kScheduleLoop = 1 };
// Device memory a fused launch reduces into instead of the workspace: block
// partials (room for max_partials doubles), a zeroed counter and the totals.
struct ScratchTarget {
  double* partials;
  size_t max_partials;
  unsigned* counter;
  double* totals;
};
// With `into`, the launch only enqueues on st and h_sums is unused; without,
// it reduces through the workspace and returns the sums in h_sums.
int launch_fused(int model, int precision, int schedule, const void* mconst,
                 const StreamDesc& sd, uint64_t n_paths, uint64_t steps, const double* strikes,
                 int n_strikes, double* h_sums, hipStream_t st, const Err& err,
                 const ScratchTarget* into = nullptr);
// The fused calls with nothing allocated: partials in the file's static
// device scratch, prices / standard errors / sums (d_sums may be null)
// written to the caller's device arrays; returns after the stream sync.
int launch_fused_noalloc(int model, int precision, int schedule, const void* mconst,
                         const StreamDesc& sd, uint64_t n_paths, uint64_t steps,
                         const double* strikes, int n_strikes, double disc, double* d_price,
                         double* d_std_err, double* d_sums, hipStream_t st, const Err& err);
// greeks_fd scenarios on shared draws: consts is an array of n (<= 8)
// GbmConst / HestonConst / MertonConst whose draw grammar is the same (same
// steps; for Merton the same lambda*dt); payoff max(sign*(S_T - strike), 0);
// h_sums[i] = sum over paths of scenario i's payoff.
int launch_fd(int model, const void* consts, int n, double strike, double sign,
              const StreamDesc& sd, uint64_t n_paths, uint64_t steps, double* h_sums,
              hipStream_t st, const Err& err);
// S_T of paths 0..n_paths-1 into d_out (device), FP64 reference arithmetic.
int launch_terminals(int model, const void* mconst, const StreamDesc& sd, uint64_t n_paths,
                     uint64_t steps, double* d_out, hipStream_t st, const Err& err);
// Fixed-order sum of partials[b][i] over b into out[i] (one block, on st).
int launch_reduce_partials(const double* part, int n_blocks, int na, double* out, hipStream_t st,
                           const Err& err);
// basket.hip: h_sums = {sum payoff, sum payoff^2, d x (dP/ds0, dP/dmu, dP/dsigma)}
int run_basket(const double* s0, const double* mu, const double* sigma, const double* L,
               const double* w, uint64_t d, double strike, double t, uint64_t steps,
               uint64_t n_paths, uint64_t seed, int precision, bool greeks, bool mfma,
               double* h_sums, hipStream_t st, const Err& err);
// batch.hip
struct BatchPlan;
int batch_plan_create(const stocha_batch_instance* in, uint64_t n, uint64_t seed, int block,
                      int paths_per_lane, int waves_per_simd, BatchPlan** out, const Err& err);
int batch_plan_run(BatchPlan* P, double* price, double* std_err, const Err& err);
void batch_plan_destroy(BatchPlan* P);
uint32_t batch_plan_tiles(const BatchPlan* P);
unsigned batch_plan_grid(const BatchPlan* P);
// lsmc.hip
int run_lsmc(const GbmConst& gc, const StreamDesc& sd, uint64_t n_paths, uint64_t steps,
             double k, double discount, int is_put, int ncols, int precision, double* price,
             double* std_err, hipStream_t st, const Err& err);
// greeks.hip
int run_greeks_pathwise(const GbmConst& gc, const StreamDesc& sd, uint64_t n_paths,
                        uint64_t steps, double s0, double r, double sigma, double t,
                        double strike, int is_call, double* delta_sum, double* vega_sum,
                        hipStream_t st, const Err& err);

}  // namespace stocha
