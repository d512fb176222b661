// streams.hip -- the random-number primitives on their own: raw PCG64DXSM
// draws and ziggurat normals of the block-split streams (pcg64.rs:23-43,
// normal.rs:18-74), one thread per path, output [path][draw] row-major.
// The steppers inline the same device functions (rng.hpp).
#include "host.hpp"
#include "drivers.hpp"

namespace stocha {

constexpr int kStreamBlock = 128;

__global__ __launch_bounds__(kStreamBlock) void stream_u64_kernel(const StreamDesc sd,
                                                                  const uint64_t path0,
                                                                  const uint64_t n_paths,
                                                                  const uint64_t n_draws,
                                                                  uint64_t* __restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * kStreamBlock + threadIdx.x;
  if (i >= n_paths) return;
  Pcg g{stream_at(&sd, path0 + i)};
  for (uint64_t j = 0; j < n_draws; ++j) out[i * n_draws + j] = g.next(sd.inc);
}

// precision 64: FP64 normals; 32: FP32 normals (stored as double).  Also
// writes the number of u64 draws each path consumed.
template <bool kFp32>
__global__ __launch_bounds__(kStreamBlock) void stream_normal_kernel(
    const StreamDesc sd, const uint64_t path0, const uint64_t n_paths, const uint64_t n_draws,
    double* __restrict__ out, uint64_t* __restrict__ consumed) {
  __shared__ ZigLds zig;
  zig_load(&zig);
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * kStreamBlock + threadIdx.x;
  if (i >= n_paths) return;
  Pcg g{stream_at(&sd, path0 + i)};
  const U128 start = g.s;
  for (uint64_t j = 0; j < n_draws; ++j)
    out[i * n_draws + j] = kFp32 ? (double)zig_normal32(g, sd.inc, &zig)
                                 : zig_normal64(g, sd.inc, &zig);
  if (consumed) {
    // distance from start: step a copy until it meets the final state
    Pcg h{start};
    uint64_t c = 0;
    while ((h.s.lo != g.s.lo || h.s.hi != g.s.hi) && c < 64 * n_draws + 64) {
      (void)h.next(sd.inc);
      ++c;
    }
    consumed[i] = c;
  }
}

// The same normals drawn on the queue schedule of the steppers (drivers.hpp:
// ZigWave's lane-mask state machine, per-lane rings, deferred slow block) --
// another schedule of the same draws, so the same values: a test of that
// schedule on chosen stream positions (e.g. the rare fast-test ties).  The
// model stores each step's normal in its row.
template <bool kFp32>
struct NormalRow {
  double* row;  // nullptr: a padding lane of the last wavefront
  uint32_t k = 0;
  __device__ __forceinline__ NormalRow(double* r) : row(r) {}
  __device__ __forceinline__ void apply(typename ZigWave<kFp32>::Real z) {
    if (row) row[k] = (double)z;
    ++k;
  }
};

template <bool kFp32>
__global__ __launch_bounds__(kStreamBlock) void stream_normal_queue_kernel(
    const StreamDesc sd, const uint64_t path0, const uint64_t n_paths, const uint32_t n_draws,
    double* __restrict__ out) {
  using Src = NormalSrc<kFp32, 1>;
  constexpr uint32_t kSlots = 16;
  __shared__ typename Src::Tables zig;
  __shared__ typename Src::Rec q[kSlots][kStreamBlock];
  zig_load(&zig);
  __syncthreads();
  const uint64_t i = (uint64_t)blockIdx.x * kStreamBlock + threadIdx.x;
  // whole wavefronts past the last path leave; padding lanes of the last one
  // replay path path0 and store nothing
  if ((uint64_t)blockIdx.x * kStreamBlock + (threadIdx.x & ~63u) >= n_paths) return;
  const bool valid = i < n_paths;
  Pcg g{stream_at(&sd, path0 + (valid ? i : 0))};
  Src src(sd);
  NormalRow<kFp32> m(valid ? out + i * n_draws : nullptr);
  queue_run<Src, kSlots, kStreamBlock>(src, m, g, sd.inc, n_draws, &zig, q);
}

}  // namespace stocha

using namespace stocha;

extern "C" int stocha_stream_u64(uint64_t seed, uint64_t block, uint64_t path0, uint64_t n_paths,
                                 uint64_t n_draws, uint64_t* out, char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  if (n_paths == 0 || n_draws == 0) return STOCHA_OK;
  std::lock_guard<std::mutex> lk(globals().mu);
  hipStream_t st = globals().stream;
  StreamDesc sd;
  make_stream_desc(seed, block, &sd);
  DevBuf d;
  const size_t bytes = sizeof(uint64_t) * n_paths * n_draws;
  if (int rc = d.alloc(bytes, err)) return rc;
  stream_u64_kernel<<<(unsigned)((n_paths + kStreamBlock - 1) / kStreamBlock), kStreamBlock, 0,
                      st>>>(sd, path0, n_paths, n_draws, d.as<uint64_t>());
  STOCHA_HIP(hipGetLastError());
  STOCHA_HIP(hipMemcpyAsync(out, d.p, bytes, hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
  return STOCHA_OK;
}

extern "C" int stocha_stream_normals(uint64_t seed, uint64_t block, uint64_t path0,
                                     uint64_t n_paths, uint64_t n_draws, int precision,
                                     double* out, uint64_t* consumed, char* errbuf,
                                     size_t errlen) {
  Err err{errbuf, errlen};
  if (precision != STOCHA_FP64 && precision != STOCHA_FP32)
    return err.set(STOCHA_EINVAL, "precision must be 32 or 64");
  if (n_paths == 0 || n_draws == 0) return STOCHA_OK;
  std::lock_guard<std::mutex> lk(globals().mu);
  hipStream_t st = globals().stream;
  StreamDesc sd;
  make_stream_desc(seed, block, &sd);
  DevBuf d, c;
  const size_t bytes = sizeof(double) * n_paths * n_draws;
  if (int rc = d.alloc(bytes, err)) return rc;
  if (int rc = c.alloc(sizeof(uint64_t) * n_paths, err)) return rc;
  const unsigned grid = (unsigned)((n_paths + kStreamBlock - 1) / kStreamBlock);
  if (precision == STOCHA_FP32)
    stream_normal_kernel<true><<<grid, kStreamBlock, 0, st>>>(sd, path0, n_paths, n_draws,
                                                              d.as<double>(), c.as<uint64_t>());
  else
    stream_normal_kernel<false><<<grid, kStreamBlock, 0, st>>>(sd, path0, n_paths, n_draws,
                                                               d.as<double>(), c.as<uint64_t>());
  STOCHA_HIP(hipGetLastError());
  STOCHA_HIP(hipMemcpyAsync(out, d.p, bytes, hipMemcpyDeviceToHost, st));
  if (consumed)
    STOCHA_HIP(hipMemcpyAsync(consumed, c.p, sizeof(uint64_t) * n_paths, hipMemcpyDeviceToHost,
                              st));
  STOCHA_HIP(hipStreamSynchronize(st));
  return STOCHA_OK;
}

extern "C" int stocha_stream_normals_queue(uint64_t seed, uint64_t block, uint64_t path0,
                                           uint64_t n_paths, uint64_t n_draws, int precision,
                                           double* out, char* errbuf, size_t errlen) {
  Err err{errbuf, errlen};
  if (precision != STOCHA_FP64 && precision != STOCHA_FP32)
    return err.set(STOCHA_EINVAL, "precision must be 32 or 64");
  if (n_draws > 0xffffffffull) return err.set(STOCHA_EINVAL, "n_draws must be below 2^32");
  if (n_paths == 0 || n_draws == 0) return STOCHA_OK;
  const uint64_t grid64 = (n_paths + kStreamBlock - 1) / kStreamBlock;
  if (grid64 > 0x7fffffffull) return err.set(STOCHA_EINVAL, "n_paths too large for one launch");
  if (n_paths > ~(uint64_t)0 / n_draws / sizeof(double))
    return err.set(STOCHA_EINVAL, "n_paths x n_draws too large");
  std::lock_guard<std::mutex> lk(globals().mu);
  hipStream_t st = globals().stream;
  StreamDesc sd;
  make_stream_desc(seed, block, &sd);
  DevBuf d;
  const size_t bytes = sizeof(double) * n_paths * n_draws;
  if (int rc = d.alloc(bytes, err)) return rc;
  const unsigned grid = (unsigned)grid64;
  if (precision == STOCHA_FP32)
    stream_normal_queue_kernel<true><<<grid, kStreamBlock, 0, st>>>(sd, path0, n_paths,
                                                                    (uint32_t)n_draws,
                                                                    d.as<double>());
  else
    stream_normal_queue_kernel<false><<<grid, kStreamBlock, 0, st>>>(sd, path0, n_paths,
                                                                     (uint32_t)n_draws,
                                                                     d.as<double>());
  STOCHA_HIP(hipGetLastError());
  STOCHA_HIP(hipMemcpyAsync(out, d.p, bytes, hipMemcpyDeviceToHost, st));
  STOCHA_HIP(hipStreamSynchronize(st));
  return STOCHA_OK;
}
